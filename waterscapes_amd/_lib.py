"""ctypes binding of libmpet_hip.so (declarations: include/mpet_hip.h).

The product path has no CPU fallback: if the library is missing or no GPU is
visible, every device entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import refel

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmpet_hip.so")

MAXF = 8
TAB_SIZE = 1860


class MeshDesc(C.Structure):
    _fields_ = [("nV", C.c_int32), ("nC", C.c_int32), ("nP2", C.c_int32), ("nP1", C.c_int32),
                ("coords", C.c_void_p), ("cells", C.c_void_p), ("cell_p2", C.c_void_p),
                ("colour_cells", C.c_void_p), ("colour_offsets", C.c_void_p),
                ("ncolours", C.c_int32)]


class Patterns(C.Structure):
    _fields_ = [("rp22", C.c_void_p), ("ci22", C.c_void_p), ("rp21", C.c_void_p),
                ("ci21", C.c_void_p), ("rp12", C.c_void_p), ("ci12", C.c_void_p),
                ("rp11", C.c_void_p), ("ci11", C.c_void_p), ("nnz22", C.c_int32),
                ("nnz21", C.c_int32), ("nnz12", C.c_int32), ("nnz11", C.c_int32)]


class Primitives(C.Structure):
    _fields_ = [("Kuu", C.c_void_p), ("Lap", C.c_void_p), ("Dv", C.c_void_p),
                ("Dt", C.c_void_p), ("M1", C.c_void_p), ("St", C.c_void_p)]


class Coef(C.Structure):
    _fields_ = [("C", C.c_double * (MAXF * MAXF)), ("stiff", C.c_double * MAXF),
                ("mu", C.c_double)]


class OpDesc(C.Structure):
    _fields_ = [("nP2", C.c_int32), ("nP1", C.c_int32), ("A", C.c_int32), ("pat", Patterns),
                ("dtype", C.c_int32), ("Kuu", C.c_void_p), ("Dv", C.c_void_p),
                ("Dt", C.c_void_p), ("M1", C.c_void_p), ("St", C.c_void_p), ("coef", Coef),
                ("dmask", C.c_void_p), ("lanes_u", C.c_int32), ("lanes_p", C.c_int32),
                ("xcd_chunk", C.c_int32)]


class CellOpDesc(C.Structure):
    _fields_ = [("nC", C.c_int32), ("nP2", C.c_int32), ("nP1", C.c_int32), ("A", C.c_int32),
                ("node2", C.c_void_p), ("node1", C.c_void_p), ("geom", C.c_void_p),
                ("cells_per_block", C.c_int32), ("npart2", C.c_int32), ("npart1", C.c_int32),
                ("bq2", C.c_void_p), ("lrp2", C.c_void_p), ("pe2", C.c_void_p),
                ("lsl2", C.c_void_p), ("bq1", C.c_void_p), ("lrp1", C.c_void_p),
                ("pe1", C.c_void_p), ("lsl1", C.c_void_p), ("inc2_rp", C.c_void_p),
                ("inc1_rp", C.c_void_p), ("rp11", C.c_void_p), ("ci11", C.c_void_p),
                ("nnz11", C.c_int32), ("M1", C.c_void_p), ("St", C.c_void_p), ("coef", Coef),
                ("dmask", C.c_void_p), ("lanes_u", C.c_int32), ("lanes_p", C.c_int32),
                ("xcd_chunk", C.c_int32), ("fp64", C.c_int32)]


class TwoFieldCoef(C.Structure):
    _fields_ = [("C", C.c_double * (MAXF * MAXF)), ("stiff", C.c_double * MAXF),
                ("alpha", C.c_double * MAXF), ("mu", C.c_double), ("lmbda", C.c_double)]


def make_two_field_coef(Cm, stiff, alpha, mu, lmbda):
    c = TwoFieldCoef()
    F = Cm.shape[0]
    for f in range(F):
        for g in range(F):
            c.C[f * MAXF + g] = float(Cm[f, g])
        c.stiff[f] = float(stiff[f])
        c.alpha[f] = float(alpha[f])
    c.mu, c.lmbda = float(mu), float(lmbda)
    return c


def make_coef(Cm, stiff, mu=0.0):
    c = Coef()
    F = Cm.shape[0]
    for f in range(F):
        for g in range(F):
            c.C[f * MAXF + g] = float(Cm[f, g])
        c.stiff[f] = float(stiff[f])
    c.mu = float(mu)
    return c


class MpetError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmpet_hip error {code}: {msg}")
        self.code = code


_lib = None
_tables_uploaded = set()


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            f"{LIBPATH} not found: build it with `python -m waterscapes_amd.build` "
            "(the MI355X backend has no CPU fallback)")
    L = C.CDLL(LIBPATH)
    P = C.c_void_p
    sigs = {
        "mpet_last_error": ([C.c_char_p, C.c_size_t], C.c_int),
        "mpet_upload_tables": ([P, C.c_int], C.c_int),
        "mpet_assemble_primitives": ([P, P, P, C.c_int, C.c_double, P, P], C.c_int),
        "mpet_assemble_primitives_batched": ([C.c_int, P, P, P, P, P, P, P], C.c_int),
        "mpet_expand_csr": ([P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, P, P, P], C.c_int),
        "mpet_tf_expand_csr": ([P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, P, P, P],
                               C.c_int),
        "mpet_assemble_p2_mass": ([P, P, P, P], C.c_int),
        "mpet_error_indicators": ([P, P, P, P, C.c_int, P, P, P, P], C.c_int),
        "mpet_error_indicators_noalloc": ([P, P, P, P, C.c_int, P, P, P], C.c_int),
        "mpet_constraint_rows": ([P, P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, C.c_int,
                                  P, P], C.c_int),
        "mpet_rhs_pressure": ([P, P, C.c_int, C.c_int, C.c_int, P, P, P, P], C.c_int),
        "mpet_cell_source": ([P, P, C.c_int, P, C.c_int, P, C.c_int, C.c_double,
                              C.c_longlong, C.c_int, C.c_int, P, P], C.c_int),
        "mpet_op_apply": ([P, P, P, P, P], C.c_int),
        "mpet_cell_geometry": ([P, P, P], C.c_int),
        "mpet_cell_geometry64": ([P, P, P], C.c_int),
        "mpet_cell_op_apply": ([P, P, P, P, P, P], C.c_int),
        "mpet_solver_set_cell_operator": ([P, P], C.c_int),
        "mpet_solver_set_cell_operator64": ([P, P], C.c_int),
        "mpet_csr_spmv": ([C.c_int, C.c_longlong, P, P, P, P, P, P, P, C.c_int, P], C.c_int),
        "mpet_soa_convert": ([C.c_int, P, C.c_int, C.c_longlong, P, P], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names every C entry point the header declares (for load tests)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "mpet_hip.h")
    src = open(hdr).read()
    return sorted(set(re.findall(r"^\s*int\s+(mpet_\w+)\s*\(", src, re.M)))


def check(rc):
    if rc != 0:
        buf = C.create_string_buffer(512)
        lib().mpet_last_error(buf, 512)
        raise MpetError(rc, buf.value.decode(errors="replace"))


def ensure_tables(device_index):
    if device_index in _tables_uploaded:
        return
    tab = refel.tables_float64()
    assert tab.size == TAB_SIZE
    check(lib().mpet_upload_tables(tab.ctypes.data_as(C.c_void_p), TAB_SIZE))
    _tables_uploaded.add(device_index)


def ptr(t):
    """Device/host address of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
