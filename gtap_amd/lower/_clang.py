"""Minimal ctypes binding of libclang's C API (the subset the lowering tool uses).

libclang ships with ROCm's LLVM (/opt/rocm/llvm/lib/libclang.so); its C headers are not
installed, so the few types and entry points are declared here from the published
clang-c/Index.h interface.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

LIBCLANG = os.environ.get("GTAP_LIBCLANG", "/opt/rocm/llvm/lib/libclang.so")

# CXCursorKind values used by the lowering
FUNCTION_DECL = 8
VAR_DECL = 9
PARM_DECL = 10
DECL_REF_EXPR = 101
MEMBER_REF_EXPR = 102
CALL_EXPR = 103
ARRAY_SUBSCRIPT_EXPR = 113
BINARY_OPERATOR = 114
COMPOUND_STMT = 202
FOR_STMT = 209
RETURN_STMT = 214
DECL_STMT = 231


class _CXString(C.Structure):
    _fields_ = [("data", C.c_void_p), ("flags", C.c_uint)]


class _CXCursor(C.Structure):
    _fields_ = [("kind", C.c_int), ("xdata", C.c_int), ("data", C.c_void_p * 3)]


class _CXSourceLocation(C.Structure):
    _fields_ = [("ptr_data", C.c_void_p * 2), ("int_data", C.c_uint)]


class _CXSourceRange(C.Structure):
    _fields_ = [("ptr_data", C.c_void_p * 2), ("begin", C.c_uint), ("end", C.c_uint)]


class _CXType(C.Structure):
    _fields_ = [("kind", C.c_int), ("data", C.c_void_p * 2)]


_VISITOR = C.CFUNCTYPE(C.c_int, _CXCursor, _CXCursor, C.py_object)
_lib: Optional[C.CDLL] = None


def _load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    lib = C.CDLL(LIBCLANG)
    sig = {
        "clang_createIndex": ([C.c_int, C.c_int], C.c_void_p),
        "clang_disposeIndex": ([C.c_void_p], None),
        "clang_parseTranslationUnit": ([C.c_void_p, C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.c_void_p,
                                        C.c_uint, C.c_uint], C.c_void_p),
        "clang_disposeTranslationUnit": ([C.c_void_p], None),
        "clang_getTranslationUnitCursor": ([C.c_void_p], _CXCursor),
        "clang_visitChildren": ([_CXCursor, _VISITOR, C.py_object], C.c_uint),
        "clang_getCursorKind": ([_CXCursor], C.c_int),
        "clang_getCursorSpelling": ([_CXCursor], _CXString),
        "clang_getCString": ([_CXString], C.c_char_p),
        "clang_disposeString": ([_CXString], None),
        "clang_getCursorExtent": ([_CXCursor], _CXSourceRange),
        "clang_getCursorLocation": ([_CXCursor], _CXSourceLocation),
        "clang_getRangeStart": ([_CXSourceRange], _CXSourceLocation),
        "clang_getRangeEnd": ([_CXSourceRange], _CXSourceLocation),
        "clang_getFileLocation": ([_CXSourceLocation, C.POINTER(C.c_void_p), C.POINTER(C.c_uint),
                                   C.POINTER(C.c_uint), C.POINTER(C.c_uint)], None),
        "clang_getFileName": ([C.c_void_p], _CXString),
        "clang_getCursorType": ([_CXCursor], _CXType),
        "clang_getCursorResultType": ([_CXCursor], _CXType),
        "clang_getTypeSpelling": ([_CXType], _CXString),
        "clang_Type_getSizeOf": ([_CXType], C.c_longlong),
        "clang_getCanonicalType": ([_CXType], _CXType),
        "clang_getArrayElementType": ([_CXType], _CXType),
        "clang_getArraySize": ([_CXType], C.c_longlong),
        "clang_getCursorReferenced": ([_CXCursor], _CXCursor),
        "clang_Cursor_isNull": ([_CXCursor], C.c_int),
        "clang_getNumDiagnostics": ([C.c_void_p], C.c_uint),
        "clang_getDiagnostic": ([C.c_void_p, C.c_uint], C.c_void_p),
        "clang_getDiagnosticSeverity": ([C.c_void_p], C.c_int),
        "clang_formatDiagnostic": ([C.c_void_p, C.c_uint], _CXString),
        "clang_disposeDiagnostic": ([C.c_void_p], None),
        "clang_getDiagnosticLocation": ([C.c_void_p], _CXSourceLocation),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _s(x: _CXString) -> str:
    lib = _load()
    v = lib.clang_getCString(x)
    out = v.decode() if v else ""
    lib.clang_disposeString(x)
    return out


@dataclass
class Node:
    """A cursor, flattened: kind, spelling, type, byte extent in the main file, children."""
    kind: int
    spelling: str
    type: str
    type_size: int
    start: int
    end: int
    file: str
    ref: Optional[tuple] = None          # (file, offset) of the referenced declaration
    ref_kind: int = 0
    elem_type: str = ""                  # array element type (VarDecl of array type)
    elem_size: int = 0
    array_len: int = -1
    result_type: str = ""                # FunctionDecl result type
    result_size: int = 0
    children: List["Node"] = field(default_factory=list)

    def walk(self):
        yield self
        for c in self.children:
            yield from c.walk()


def _loc(loc) -> tuple:
    lib = _load()
    f = C.c_void_p()
    off = C.c_uint()
    lib.clang_getFileLocation(loc, C.byref(f), None, None, C.byref(off))
    name = _s(lib.clang_getFileName(f)) if f.value else ""
    return name, off.value


def parse(path: str, args: List[str]) -> tuple[List[Node], List[tuple]]:
    """Parses `path`; returns the top-level declarations of that file and the error
    diagnostics as (file, offset, message)."""
    lib = _load()
    idx = lib.clang_createIndex(0, 0)
    argv = (C.c_char_p * len(args))(*[a.encode() for a in args])
    tu = lib.clang_parseTranslationUnit(idx, path.encode(), argv, len(args), None, 0, 0)
    if not tu:
        lib.clang_disposeIndex(idx)
        raise RuntimeError(f"libclang could not parse {path}")
    errors = []
    for i in range(lib.clang_getNumDiagnostics(tu)):
        d = lib.clang_getDiagnostic(tu, i)
        if lib.clang_getDiagnosticSeverity(d) >= 3:  # error, fatal
            f, off = _loc(lib.clang_getDiagnosticLocation(d))
            errors.append((f, off, _s(lib.clang_formatDiagnostic(d, 0))))
        lib.clang_disposeDiagnostic(d)
    main = os.path.realpath(path)

    def build(c) -> Node:
        kind = lib.clang_getCursorKind(c)
        ext = lib.clang_getCursorExtent(c)
        f0, a = _loc(lib.clang_getRangeStart(ext))
        _, b = _loc(lib.clang_getRangeEnd(ext))
        t = lib.clang_getCursorType(c)
        n = Node(kind=kind, spelling=_s(lib.clang_getCursorSpelling(c)), type=_s(lib.clang_getTypeSpelling(t)),
                 type_size=int(lib.clang_Type_getSizeOf(t)), start=a, end=b, file=f0)
        if kind in (VAR_DECL, PARM_DECL):
            ct = lib.clang_getCanonicalType(t)
            n.array_len = int(lib.clang_getArraySize(ct))
            if n.array_len >= 0:
                et = lib.clang_getArrayElementType(ct)
                n.elem_type = _s(lib.clang_getTypeSpelling(et))
                n.elem_size = int(lib.clang_Type_getSizeOf(et))
        if kind == FUNCTION_DECL:
            rt = lib.clang_getCursorResultType(c)
            n.result_type = _s(lib.clang_getTypeSpelling(rt))
            n.result_size = int(lib.clang_Type_getSizeOf(rt))
        if kind in (DECL_REF_EXPR, MEMBER_REF_EXPR, CALL_EXPR):
            r = lib.clang_getCursorReferenced(c)
            if not lib.clang_Cursor_isNull(r):
                n.ref = _loc(lib.clang_getCursorLocation(r))
                n.ref_kind = lib.clang_getCursorKind(r)
        kids: List[Node] = []

        def visit(ch, parent, data):
            kids.append(build(ch))
            return 1  # CXChildVisit_Continue

        cb = _VISITOR(visit)
        lib.clang_visitChildren(c, cb, None)
        n.children = kids
        return n

    tops: List[Node] = []

    def visit_top(ch, parent, data):
        f, _ = _loc(lib.clang_getCursorLocation(ch))
        if f and os.path.realpath(f) == main:
            tops.append(build(ch))
        return 1

    cb_top = _VISITOR(visit_top)
    lib.clang_visitChildren(lib.clang_getTranslationUnitCursor(tu), cb_top, None)
    lib.clang_disposeTranslationUnit(tu)
    lib.clang_disposeIndex(idx)
    return tops, errors
