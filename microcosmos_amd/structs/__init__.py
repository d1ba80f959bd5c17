"""State pytrees of the reference (structs/nodes.py, edges.py, fields.py) as
dataclasses of device tensors."""
from .nodes import Nodes
from .edges import Edges
from .fields import Fields

__all__ = ["Nodes", "Edges", "Fields"]
