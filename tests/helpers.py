"""Shared helpers of the test-suite: oracle RVEs built from the same specs
the product consumes, and the error measures used in the parity bars."""
import ctypes as C

import numpy as np

from madmax_amd import spec_fibers
from oracle import pyoracle as po


def oracle_rve(spec):
    """Oracle RVE for a product RveSpec (same mesh, fibers, materials)."""
    m = po.mesh(3, spec.nex, spec.ney, spec.nez, spec.lx, spec.ly, spec.lz)
    et = po.phase_map(m, spec_fibers(spec)) if spec.nfib else np.zeros(m.nelm, dtype=np.int32)
    return po.Rve(3, (spec.nex, spec.ney, spec.nez), et, [spec.young[0], spec.young[1]],
                  [spec.poisson[0], spec.poisson[1]], (spec.lx, spec.ly, spec.lz)), et


def normwise_rel(a, b):
    """max |a - b| / max |b|: the relative error of a stress vector or a
    tangent matrix (entry-wise ratios are meaningless for entries that are
    zero up to round-off, such as the normal/shear couplings)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


class DevBuf:
    """Device memory of the caller, from the HIP runtime the library itself
    links (libamdhip64.so.7), so that no second runtime enters the process."""

    def __init__(self, host):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = len(host)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * self.n)) == 0
        self.put(host)

    def put(self, host):
        h = np.ascontiguousarray(host, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0

    def get(self):
        h = np.zeros(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return h

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
