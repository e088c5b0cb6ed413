"""sem_apply_K_noalloc (the allocation-free operator entry point) through the C-ABI against
sem_apply_K, which shares its kernels: bitwise equal in both forms, on two calls in a row
whose second input is smaller (a smaller mesh uploaded into the same context), on a mesh
with a node that belongs to no element, and a mesh beyond the fixed scratch is refused."""
import ctypes

import numpy as np
import pytest

import semhip
from semhip import engine as E

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory from the HIP runtime that libsemhip.so itself resolved: the functions
    are looked up through the library's handle, whose search covers its dependencies.  A
    process can hold a second copy of the runtime (torch's), which does not see the device
    the library already holds."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = semhip._lib.load()
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.n = host.size
        self.ptr = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), 8 * self.n) == 0
        self.put(host)

    def put(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.size == self.n
        assert self.hip().hipMemcpy(self.ptr, host.ctypes.data, 8 * self.n, 1) == 0  # host to device

    def get(self):
        out = np.empty(self.n)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        self.hip().hipFree(self.ptr)


def _noalloc(s, u, precision):
    u_d, f_d = DeviceBuffer(u), DeviceBuffer(np.full(u.size, np.nan))
    try:
        s.apply_K_noalloc(u_d.ptr.value, f_d.ptr.value, precision)
        return f_d.get()
    finally:
        u_d.free()
        f_d.free()


def _shell(s, n, order, extra_node=False):
    q, p = semhip.rect_quads(0.1, 0.08, n, n - 1)
    nodes, coords = semhip.quad2spectral(q, p, order)
    if extra_node:  # a node that belongs to no element
        coords = np.vstack([coords, [[0.3, 0.3, 0.0]]])
    lam = semhip.laminate_shell(np.full(8, 0.25e-3), [0, 90, 0, 90, 90, 0, 90, 0], 1570.0,
                                semhip.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3))
    s.upload_mesh(semhip.SEM_SHELL5, nodes, coords, order + 1)
    s.set_material(lam)


def _solid(s, n, order):
    q, p = semhip.rect_quads(0.02, 0.02, n, n)
    n2, c2 = semhip.quad2spectral(q, p, order)
    nodes, coords = semhip.extrude(n2, c2, order + 1, order + 1, [1e-3, 1e-3])
    q3 = semhip.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    lam = semhip.laminate_solid_smeared(np.full(8, 0.25e-3), [0, 90, 0, 90, 90, 0, 90, 0], 1570.0, q3)
    s.upload_mesh(semhip.SEM_SOLID3, nodes, coords, order + 1, order + 1)
    s.set_material(lam)


@pytest.mark.parametrize("precision", ["fp64", "fp32_split"])
def test_noalloc_equals_apply_K_on_two_calls(precision):
    rng = np.random.default_rng(7)
    with semhip.Solver(0) as s:
        s.set_kernel(1, 8)
        for build, big, small in ((_shell, (12, 5), (5, 4)), (_solid, (6, 5), (3, 4))):
            for n, order in (big, small):  # the second call's input is smaller
                print(f"noalloc {build.__name__} n={n} order={order} {precision}", flush=True)
                build(s, n, order)
                u = rng.standard_normal(s.dof)
                ref = s.apply_K(u, precision)
                got = _noalloc(s, u, precision)
                assert np.array_equal(got, ref), (build.__name__, n, order, np.abs(got - ref).max())
                assert np.abs(ref).max() > 0


def test_noalloc_zero_on_a_node_of_no_element():
    rng = np.random.default_rng(8)
    with semhip.Solver(0) as s:
        s.set_kernel(1, 8)
        _shell(s, 6, 5, extra_node=True)
        u = rng.standard_normal(s.dof)
        ref = s.apply_K(u)
        got = _noalloc(s, u, "fp64")
        assert np.array_equal(got, ref)
        assert (got[-5:] == 0.0).all() and (ref[-5:] == 0.0).all()


def test_noalloc_refuses_a_mesh_beyond_its_scratch():
    with semhip.Solver(0) as s:
        s.set_kernel(1, 8)
        _shell(s, 700, 2)  # 1401 x 1400 nodes, 9.8M DOF: 2 dof doubles exceed 128 MiB
        u_d, f_d = DeviceBuffer(np.zeros(s.dof)), DeviceBuffer(np.zeros(s.dof))
        try:
            with pytest.raises(semhip.SemError) as e:
                s.apply_K_noalloc(u_d.ptr.value, f_d.ptr.value)
        finally:
            u_d.free()
            f_d.free()
        assert e.value.code == semhip._lib.SEM_E_ARG and "scratch" in str(e.value)
        ref = s.apply_K(np.ones(s.dof))  # the context stays usable
        assert np.isfinite(ref).all()
