"""Continuous-batching serving of streaming sosfilt on MI355X."""

from .engine import HipExecutor, Request, ServeStats, ServingEngine, StreamTable, run_plain

__all__ = ["HipExecutor", "Request", "ServeStats", "ServingEngine", "StreamTable", "run_plain"]
