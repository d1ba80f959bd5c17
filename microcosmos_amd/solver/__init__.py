"""Solver layer: configuration, rod constraints, fluid coupling."""
