"""mdt_compute_emt_noalloc: the allocation-free variant of the compute_emt
drop-in (device arrays in and out, the universe staged in the library's
per-GPU scratch array) against mdt_compute_emt on the same universes.

Two calls in a row, the second on a smaller universe (the scratch array
keeps the first call's contents), each to the parity bar of the other EMT
tests (energies 1e-9 relative, forces 1e-9 eV/A, densities 1e-11) and
with the drop-in's semantics: epot and f accumulate onto what the caller
passed, dens is set for the EMT atoms only, status carries the call's
bits.  Universes beyond the scratch limits are refused.  The test's own
device buffers come from the HIP runtime; the entry point allocates none.
"""
import ctypes as C

import numpy as np
import pytest

from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    mdt.lib()  # loads libamdhip64 with the library
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _last_error():
    buf = C.create_string_buffer(2048)
    mdt.lib().mdt_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


class _Dev:
    """A device copy of a numpy array (the test's buffers, not the library's)."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(1, self.a.nbytes)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, _H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(ctx, hip, s, e0, f0, d0):
    N, P = s.natoms, s.nbeads
    bufs = [_Dev(hip, a) for a in (np.asarray(s.idx + 1, dtype=np.int32),
                                   np.asarray(s.simbox, dtype=np.float64).T.copy(),
                                   np.linalg.inv(s.simbox).T.copy(), np.asarray(s.r, dtype=np.float64),
                                   e0, f0, d0, np.array([-1], dtype=np.int32))]
    idx, sb, isb, r, ep, f, d, st = bufs
    rc = mdt.lib().mdt_compute_emt_noalloc(ctx.h, N, P, idx.p, sb.p, isb.p, r.p, ep.p, f.p, d.p, st.p,
                                           mdt.ENERGY_AND_FORCE, None)
    assert rc == 0, _last_error()
    assert hip.hipDeviceSynchronize() == 0  # the call returns before its kernels finish
    out = ep.get(), f.get(), d.get(), int(st.get()[0])
    for b in bufs:
        b.free()
    return out


def _check(got, ref, tag):
    (e, f, d, status), (e_r, f_r, d_r) = got, ref
    assert status == 0, tag
    assert np.abs(e - e_r).max() <= 1e-9 * max(1.0, np.abs(e_r).max()), tag
    assert np.abs(f - f_r).max() <= 1e-9, tag
    assert np.abs(d - d_r).max() <= 1e-11, tag


def test_noalloc_equals_compute_emt_two_calls(gpu_ctx, pes_hau):
    hip = _hip()
    g = np.random.default_rng(707)
    big = geometry.h_on_slab("Au", n=6, nlayers=4, rng=g, nbeads=4, z0=1.5)
    small = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, nbeads=1, z0=2.2)
    for s in (big, small):
        s.r += g.normal(0, 0.05, s.r.shape)
    cols, pesm = pesfile.emt_tables(pes_hau, big.type_names)
    gpu_ctx.set_emt(cols, pesm)
    for k, s in enumerate((big, small)):
        N, P = s.natoms, s.nbeads
        # accumulation: start from nonzero epot and f, as after another PES
        e0 = g.normal(0, 1, P)
        f0 = g.normal(0, 1, (N, P, 3))
        d0 = np.full((N, P), -7.0)
        ref = gpu_ctx.compute_emt(s, epot=e0.copy(), f=f0.copy(), dens=d0.copy())
        got = _noalloc(gpu_ctx, hip, s, e0, f0, d0)
        _check(got, ref, "call %d (N=%d, P=%d)" % (k, N, P))


def test_noalloc_limits_refused(gpu_ctx, pes_hau):
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(1))
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    L = mdt.lib()
    nul = C.c_void_p(8)  # never dereferenced: the sizes are refused first
    for N, P in ((2049, 1), (100, 65), (1024, 17)):
        rc = L.mdt_compute_emt_noalloc(gpu_ctx.h, N, P, nul, nul, nul, nul, nul, nul, nul, None,
                                       mdt.ENERGY_AND_FORCE, None)
        assert rc == -1  # MDT_EINVAL
        assert "16384" in _last_error()
