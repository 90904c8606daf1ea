"""Continuous-batching serving of streaming sosfilt on MI355X."""

from .engine import (HipExecutor, PersistentExecutor, Request, ServeStats, ServingEngine,
                     StreamTable, run_plain)

__all__ = ["HipExecutor", "PersistentExecutor", "Request", "ServeStats", "ServingEngine",
           "StreamTable", "run_plain"]
