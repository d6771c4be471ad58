"""Result/Error surface (reference: include/cusbf/error.hpp).

The reference returns Result<T> (an expected<T, Error>) whose Error carries one
of five categories with pinned indices (error.hpp:67-73, 139-143). In Python a
failing call raises CusbfError; `category` is that ErrorCategory and `message`
the Error::message() text. `Result` is provided for callers that prefer
values to exceptions (`Result.of(fn, *args)`).
"""
from __future__ import annotations

import enum
from dataclasses import dataclass
from typing import Any, Callable, Generic, Optional, TypeVar


class ErrorCategory(enum.IntEnum):
    """error.hpp:67-73; the integer values are the variant indices."""

    cuda = 0  # device runtime errors (HIP here)
    io = 1
    fastx_parse = 2
    invalid_argument = 3
    resource = 4


@dataclass(frozen=True)
class SourceLocation:
    """error.hpp:20-41: file, 1-based line and column of a FASTX parse error."""

    file: str
    line: int
    column: int

    def to_string(self) -> str:
        return f"{self.file}:{self.line}:{self.column}"


class CusbfError(RuntimeError):
    def __init__(self, category: ErrorCategory, message: str,
                 location: Optional[SourceLocation] = None):
        super().__init__(message)
        self.category = ErrorCategory(category)
        self.message = message
        self.location = location

    def __repr__(self) -> str:
        return f"CusbfError({self.category.name}, {self.message!r})"


T = TypeVar("T")


@dataclass
class Result(Generic[T]):
    """Value-or-error holder mirroring Result<T> (error.hpp:151-175)."""

    _value: Optional[T] = None
    _error: Optional[CusbfError] = None

    def __bool__(self) -> bool:
        return self._error is None

    def value(self) -> T:
        if self._error is not None:
            raise self._error
        return self._value  # type: ignore[return-value]

    def error(self) -> CusbfError:
        if self._error is None:
            raise ValueError("Result holds a value")
        return self._error

    @staticmethod
    def of(fn: Callable[..., Any], *args, **kwargs) -> "Result":
        try:
            return Result(fn(*args, **kwargs), None)
        except CusbfError as e:  # only library errors become values
            return Result(None, e)


def check(lib, status: int) -> None:
    """Raise the calling thread's last library error when status != 0."""
    if status == 0:
        return
    import ctypes

    buf = ctypes.create_string_buffer(4096)
    lib.cusbf_last_error(buf, len(buf))
    category = ErrorCategory(status - 1)
    location = None
    if category == ErrorCategory.fastx_parse:
        fbuf = ctypes.create_string_buffer(4096)
        line, col = ctypes.c_uint32(), ctypes.c_uint32()
        lib.cusbf_last_error_location(fbuf, len(fbuf), ctypes.byref(line), ctypes.byref(col))
        location = SourceLocation(fbuf.value.decode(errors="replace"), line.value, col.value)
    raise CusbfError(category, buf.value.decode(errors="replace"), location)
