"""ctypes binding of libgenesis_rigid.so (the C-ABI declared in include/genesis_rigid.h).

The ctypes mirrors of gsr_model / gsr_state are generated from the header itself, so the field order cannot drift
from the C declaration; the library's own sizeof() is checked at load time.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "genesis_rigid.h")
LIB_PATH = os.path.join(HERE, "_build", "libgenesis_rigid.so")

EXPORTS = ["gsr_abi_version", "gsr_model_size", "gsr_state_size", "gsr_group_size", "gsr_batch_info_size",
           "gsr_gjk_workspace_words",
           "gsr_step", "gsr_noalloc_scratch_bytes", "gsr_step_noalloc", "gsr_forward_kinematics", "gsr_reset_envs", "gsr_get_errno",
           "gsr_get_contacts", "gsr_links_contact_force", "gsr_get_equalities", "gsr_run_stage", "gsr_launch_kernel", "gsr_batch_table_bytes",
           "gsr_batch_upload", "gsr_step_fused", "gsr_fused_kernel_count", "gsr_launch_fused_kernel",
           "gsr_graph_capture", "gsr_graph_launch", "gsr_graph_destroy",
           "gsr_serve_host_alloc", "gsr_serve_host_free", "gsr_serve_state_bytes", "gsr_serve_table_bytes",
           "gsr_serve_clock_khz", "gsr_serve_init", "gsr_serve_tick", "gsr_serve_graph_capture"]

_CTYPE = {"int32_t": ctypes.c_int32, "uint32_t": ctypes.c_uint32, "float": ctypes.c_float,
          "int64_t": ctypes.c_int64, "uint64_t": ctypes.c_uint64, "double": ctypes.c_double}


def _defines():
    src = open(HEADER).read()
    return {k: int(v) for k, v in re.findall(r"#define (\w+) (\d+)\b", src)}


def _parse_struct(name: str):
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    consts = _defines()
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, flags=re.S)
    body = m.group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = decl.replace("const ", "")
        base, rest = decl.split(None, 1)
        for var in rest.split(","):
            var = var.strip()
            if var.startswith("*"):
                fields.append((var[1:].strip(), ctypes.c_void_p))
            elif "[" in var:
                vname, n = re.match(r"(\w+)\[(\w+)\]", var).groups()
                n = int(n) if n.isdigit() else consts[n]
                fields.append((vname, (_CTYPE[base] if base in _CTYPE else _STRUCTS[base]) * n))
            else:
                fields.append((var, _CTYPE[base] if base in _CTYPE else _STRUCTS[base]))
    return fields


_STRUCTS = {}


class GsrModel(ctypes.Structure):
    _fields_ = _parse_struct("gsr_model")


class GsrState(ctypes.Structure):
    _fields_ = _parse_struct("gsr_state")


_STRUCTS.update(gsr_model=GsrModel, gsr_state=GsrState)


class GsrGroup(ctypes.Structure):
    _fields_ = _parse_struct("gsr_group")


class GsrBatchInfo(ctypes.Structure):
    _fields_ = _parse_struct("gsr_batch_info")


class GsrServePool(ctypes.Structure):
    _fields_ = _parse_struct("gsr_serve_pool")


class GsrServeRequest(ctypes.Structure):
    _fields_ = _parse_struct("gsr_serve_request")


class GsrServeDone(ctypes.Structure):
    _fields_ = _parse_struct("gsr_serve_done")


_STRUCTS.update(gsr_serve_request=GsrServeRequest, gsr_serve_done=GsrServeDone)


class GsrServeHost(ctypes.Structure):
    _fields_ = _parse_struct("gsr_serve_host")


MODEL_POINTERS = [n for n, t in GsrModel._fields_ if t is ctypes.c_void_p]
STATE_POINTERS = [n for n, t in GsrState._fields_ if t is ctypes.c_void_p]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load the HIP library. Raises LibraryMissing if it has not been built: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} not found; run __graft_entry__.build() (hipcc --offload-arch=gfx950) first")
    L = ctypes.CDLL(path)
    for fn in EXPORTS:
        getattr(L, fn)  # raises AttributeError if a declared symbol is missing
    L.gsr_abi_version.restype = ctypes.c_int
    L.gsr_model_size.restype = ctypes.c_int
    L.gsr_state_size.restype = ctypes.c_int
    P = ctypes.POINTER
    L.gsr_step.argtypes = [P(GsrModel), P(GsrState), ctypes.c_int32, ctypes.c_void_p]
    L.gsr_noalloc_scratch_bytes.restype = ctypes.c_int64
    L.gsr_noalloc_scratch_bytes.argtypes = [P(GsrModel), ctypes.c_int32]
    L.gsr_step_noalloc.argtypes = [P(GsrModel), P(GsrState), ctypes.c_int32, ctypes.c_void_p]
    L.gsr_forward_kinematics.argtypes = [P(GsrModel), P(GsrState), ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    L.gsr_reset_envs.argtypes = [P(GsrModel), P(GsrState), ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p]
    L.gsr_get_errno.argtypes = [P(GsrState), ctypes.c_void_p, ctypes.c_void_p]
    L.gsr_get_contacts.argtypes = [P(GsrModel), P(GsrState), ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.gsr_links_contact_force.argtypes = [P(GsrModel), P(GsrState), ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                          ctypes.c_void_p]
    L.gsr_get_equalities.argtypes = [P(GsrModel), P(GsrState), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.gsr_run_stage.argtypes = [P(GsrModel), P(GsrState), ctypes.c_int32, ctypes.c_void_p]
    L.gsr_launch_kernel.argtypes = [P(GsrModel), P(GsrState), ctypes.c_int32, ctypes.c_void_p]
    L.gsr_batch_table_bytes.restype = ctypes.c_int64
    L.gsr_batch_table_bytes.argtypes = [ctypes.c_int32]
    L.gsr_batch_upload.argtypes = [P(GsrGroup), ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, P(GsrBatchInfo)]
    L.gsr_step_fused.argtypes = [ctypes.c_void_p, P(GsrBatchInfo), ctypes.c_int32, ctypes.c_void_p]
    L.gsr_fused_kernel_count.argtypes = [P(GsrBatchInfo)]
    L.gsr_launch_fused_kernel.argtypes = [ctypes.c_void_p, P(GsrBatchInfo), ctypes.c_int32, ctypes.c_void_p]
    L.gsr_graph_capture.argtypes = [ctypes.c_void_p, P(GsrBatchInfo), ctypes.c_int32, ctypes.c_void_p,
                                    P(ctypes.c_void_p)]
    L.gsr_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.gsr_graph_destroy.argtypes = [ctypes.c_void_p]
    L.gsr_serve_host_alloc.argtypes = [P(P(GsrServeHost))]
    L.gsr_serve_host_free.argtypes = [P(GsrServeHost)]
    L.gsr_serve_state_bytes.restype = ctypes.c_int64
    L.gsr_serve_state_bytes.argtypes = [ctypes.c_int32]
    L.gsr_serve_table_bytes.restype = ctypes.c_int64
    L.gsr_serve_table_bytes.argtypes = [ctypes.c_int32]
    L.gsr_serve_init.argtypes = [P(GsrServePool), ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                 ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, P(GsrBatchInfo), ctypes.c_void_p]
    L.gsr_serve_tick.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(GsrBatchInfo), P(GsrServeHost), ctypes.c_void_p]
    L.gsr_serve_graph_capture.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(GsrBatchInfo), P(GsrServeHost),
                                          ctypes.c_void_p, P(ctypes.c_void_p)]
    if (L.gsr_model_size() != ctypes.sizeof(GsrModel) or L.gsr_state_size() != ctypes.sizeof(GsrState)
            or L.gsr_group_size() != ctypes.sizeof(GsrGroup)
            or L.gsr_batch_info_size() != ctypes.sizeof(GsrBatchInfo)):
        raise RuntimeError("gsr_model/gsr_state layout mismatch between include/genesis_rigid.h and the built library")
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")
