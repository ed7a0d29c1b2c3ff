"""pdgpu_nlist_export_noalloc (the export staged in the library's static
device scratch array, no allocation) against pdgpu_nlist_export, through
the C ABI: integer output, so bit for bit. Two calls in a row on one
context, the second on a smaller instance, so the second reads an array
the first left filled; both lists (shadow on and off)."""
import ctypes as C

import numpy as np
import pytest

import pdgpu
from pdgpu import synth, workloads as W

pytestmark = pytest.mark.gpu

lib = pdgpu.lib
lib.pdgpu_nlist_export_noalloc.restype = C.c_int
lib.pdgpu_nlist_export_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                           C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_int64)]


def export_noalloc(ctx, iid, n, shadow):
    rs = np.zeros(n + 1, np.int32)
    cnt = C.c_int64()
    rc = lib.pdgpu_nlist_export_noalloc(ctx.h, iid, shadow, rs.ctypes.data_as(C.POINTER(C.c_int32)), None, 0,
                                        C.byref(cnt))
    assert rc == pdgpu.PDGPU_ERR_NOMEM and cnt.value > 0  # capacity 0: size query
    ent = np.zeros(cnt.value, np.uint32)
    rc = lib.pdgpu_nlist_export_noalloc(ctx.h, iid, shadow, rs.ctypes.data_as(C.POINTER(C.c_int32)),
                                        ent.ctypes.data_as(C.POINTER(C.c_uint32)), len(ent), C.byref(cnt))
    assert rc == pdgpu.PDGPU_OK
    return rs, ent


@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_noalloc_export_matches_export(precision):
    big = synth.peptide(seed=31, residues=120, mode="basins")
    small = synth.peptide(seed=32, residues=14)
    with pdgpu.Context(W.options("single", precision)) as ctx:
        ib = ctx.add_system(big)
        iS = ctx.add_system(small)
        # no list yet: refused rather than built (building allocates)
        rs = np.zeros(small["n"] + 1, np.int32)
        cnt = C.c_int64()
        assert lib.pdgpu_nlist_export_noalloc(ctx.h, iS, 1, rs.ctypes.data_as(C.POINTER(C.c_int32)), None, 0,
                                              C.byref(cnt)) == pdgpu.PDGPU_ERR_INVALID
        ctx.nlist_build(pdgpu.PDGPU_ALL)
        for shadow in (1, 0):
            for iid, m in ((ib, big), (iS, small)):  # the second call is the smaller one
                rs_a, ent_a = ctx.nlist_export(iid, shadow=bool(shadow))
                rs_b, ent_b = export_noalloc(ctx, iid, m["n"], shadow)
                assert np.array_equal(rs_a, rs_b)
                assert np.array_equal(ent_a, ent_b)
