"""edgpu_solve_poisson_fftx_noalloc, the FFT_X solve whose work arrays live in a static device
buffer, against edgpu_solve_poisson_fftx on the same charge densities: two calls in a row, the
second on a smaller grid, each equal to the context entry point's potential bit for bit (the
same kernels and tables); sizes beyond the buffer are refused."""
import ctypes as C

import numpy as np
import pytest

from edgpu import config

pytestmark = pytest.mark.gpu

OBJ_ABOVE, OBJ_BELOW = 1, 3


class _Hip:
    """Device buffers for the test through the HIP runtime the library itself links."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def get(self, src, n):
        out = np.empty(n, np.float64)
        assert self.h.hipMemcpy(out.ctypes.data, src, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self, p):
        self.h.hipFree(p)


def _case(scale, below, above, seed):
    prob = config.workload(2, scale)
    prob.object_phi[OBJ_BELOW], prob.object_phi[OBJ_ABOVE] = below, above
    rng = np.random.default_rng(seed)
    n = prob.nx * prob.ny
    rho_e = 200.0 + rng.standard_normal(n)
    rho_i = rho_e * (1.0 + 0.01 * rng.standard_normal(n))
    return prob, rho_e, rho_i


def _context_solve(edgpu, prob, rho_e, rho_i):
    rt = edgpu.Runtime(prob)
    try:
        rt.set_grid("rho_e", rho_e)
        rt.set_grid("rho_i", rho_i)
        return rt.solve_poisson_fftx()
    finally:
        rt.close()


def test_noalloc_fftx_matches_the_context_entry(edgpu):
    L = edgpu.lib()
    hip = _Hip()
    cases = [_case(8, -0.4, 0.7, 1), _case(16, 0.2, -0.1, 2)]  # 256 x 128 cells, then 128 x 64
    refs = [_context_solve(edgpu, *c) for c in cases]
    nmax = cases[0][0].nx * cases[0][0].ny
    bufs = [hip.alloc(8 * nmax) for _ in range(3)]
    try:
        got = []
        # two calls in a row, no context, the second on the smaller grid
        for prob, rho_e, rho_i in cases:
            print("case no-alloc FFT_X %dx%d" % (prob.gmi - 1, prob.gmj), flush=True)
            hip.put(bufs[0], rho_e)
            hip.put(bufs[1], rho_i)
            rc = L.edgpu_solve_poisson_fftx_noalloc(prob.gmi - 1, prob.gmj, prob.N_of_particles_cell,
                                                    prob.object_phi[OBJ_BELOW], prob.object_phi[OBJ_ABOVE],
                                                    bufs[0], bufs[1], bufs[2], None)
            assert rc == 0, L.edgpu_last_error()
            hip.sync()
            got.append(hip.get(bufs[2], prob.nx * prob.ny))
        for (prob, _, _), ref, phi in zip(cases, refs, got):
            assert np.array_equal(phi, ref), "max diff %g" % np.max(np.abs(phi - ref))
            P = phi.reshape(prob.ny, prob.nx)
            assert abs(P[0, 1] - prob.object_phi[OBJ_BELOW]) < 1e-12 * max(1.0, np.abs(P).max())
        # beyond the static buffer: refused before any work is queued
        for N, gmj in ((4096, 64), (64, 2048), (96, 64)):
            rc = L.edgpu_solve_poisson_fftx_noalloc(N, gmj, 200, 0.0, 0.0, bufs[0], bufs[1], bufs[2], None)
            assert rc == 1 and b"static work buffer" in L.edgpu_last_error()
        nb = C.c_int64()
        assert L.edgpu_noalloc_scratch_bytes(C.byref(nb)) == 0 and 25e6 < nb.value < 26e6
    finally:
        for b in bufs:
            hip.free(b)
