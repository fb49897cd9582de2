"""paintop's error model (paintop-ir/src/error.rs:23-115, 210-246) for the hip backend.

An error carries a class, a stable code and a message; the class fixes the
process exit code exactly as ``ErrorClass::exit_code`` does.
"""
from __future__ import annotations

from enum import Enum


class ErrorClass(str, Enum):
    """paintop_ir::ErrorClass (error.rs:23-37)."""

    PARSE = "parse"
    SCHEMA = "schema"
    REFERENCE = "reference"
    TYPE = "type"
    SEMANTIC = "semantic"
    POLICY = "policy"
    EXECUTION = "execution"
    ASSERTION = "assertion"
    CONFORMANCE = "conformance"
    MODEL = "model"
    ASSET = "asset"
    EXPORT = "export"

    def exit_code(self) -> int:
        """error.rs:103-114."""
        return {
            "parse": 2, "schema": 2, "reference": 2,
            "type": 3, "semantic": 3,
            "policy": 4,
            "execution": 5,
            "assertion": 6,
            "conformance": 7,
            "model": 8,
            "asset": 9, "export": 9,
        }[self.value]


class PaintopError(Exception):
    """paintop_ir::Error: class + stable code + message (+ optional context)."""

    def __init__(self, error_class: ErrorClass, code: str, message: str, actual: str | None = None):
        super().__init__(f"[{error_class.value}] {code}: {message}")
        self.error_class = ErrorClass(error_class)
        self.code = code
        self.message = message
        self.actual = actual

    @property
    def exit_code(self) -> int:
        return self.error_class.exit_code()


# GPU-backend codes, as the wgpu backend template names them
# (paintop-wgpu/src/gpu/error.rs:8-17, 53-66).
E_GPU_UNAVAILABLE = "E_GPU_UNAVAILABLE"  # class Policy
E_GPU_DISPATCH_INVALID = "E_GPU_DISPATCH_INVALID"  # class Execution
# executor/dispatch.rs: a required backend that is not declared and registered
E_BACKEND_UNSUPPORTED = "E_BACKEND_UNSUPPORTED"  # class Policy
E_OUTPUT_NOT_PRODUCED = "E_OUTPUT_NOT_PRODUCED"  # class Execution (executor/mod.rs:326-337)
