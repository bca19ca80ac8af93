"""iso_extract_mc_noalloc: the single-pass extraction that allocates nothing.

Its output must equal iso_extract_mc's on the same bound grid, byte for
byte, on consecutive calls (the second with a smaller lattice, so that the
shared scratch holds the first call's words); free device memory must not
change across the calls; lattices beyond its scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from isosurface_amd import _abi as A

pytestmark = pytest.mark.gpu


def blob_grid(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.zeros((nz, ny, nx), np.float64)
    for _ in range(6):
        k = rng.normal(0, 0.3, 3)
        f += np.sin(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    return np.ascontiguousarray((f - 0.3).astype(np.float32))


def desc_of(nx, ny, nz):
    d = A.IsoGridDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    for i in range(3):
        d.origin[i] = 0.0
        d.spacing[i] = 1.0 / (nx - 1)
    d.source = A.ISO_SOURCE_GRID
    return d


def free_device_bytes() -> int:
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return int(free.value)


class DeviceOut:
    def __init__(self, ctx, nv, ni):
        self.ctx, self.nv, self.ni = ctx, nv, ni
        self.v, self.i = C.c_void_p(), C.c_void_p()
        ctx.check(ctx.lib.iso_device_alloc(ctx.handle, max(12 * nv, 12), C.byref(self.v)), "alloc")
        ctx.check(ctx.lib.iso_device_alloc(ctx.handle, max(4 * ni, 4), C.byref(self.i)), "alloc")

    def out(self):
        return A.IsoMeshOut(self.v.value, self.i.value, self.nv, self.ni, 0, 0,
                            A.ISO_MEMORY_DEVICE, 0)

    def fetch(self, out):
        nv, ni = int(out.vertex_count), int(out.index_count)
        v = np.empty((nv, 3), np.float32)
        i = np.empty(ni, np.uint32)
        self.ctx.check(self.ctx.lib.iso_memcpy_to_host(self.ctx.handle, v.ctypes.data, self.v,
                                                       v.nbytes), "copy")
        self.ctx.check(self.ctx.lib.iso_memcpy_to_host(self.ctx.handle, i.ctypes.data, self.i,
                                                       i.nbytes), "copy")
        return v, i

    def free(self):
        self.ctx.lib.iso_device_free(self.ctx.handle, self.v)
        self.ctx.lib.iso_device_free(self.ctx.handle, self.i)


def test_noalloc_matches_extract_mc_on_two_calls(ctx):
    lib = ctx.lib
    for shape, seed in [((257, 257, 258), 1), ((97, 70, 41), 2)]:  # second call smaller
        nx, ny, nz = shape
        ctx.upload_grid(blob_grid(nx, ny, nz, seed))
        d = desc_of(nx, ny, nz)
        ref = ctx.extract_mc(d)
        nv, ni = len(ref.vertices), len(ref.indices)
        buf = DeviceOut(ctx, nv, ni)
        try:
            if seed == 1:  # the process's first launch of the kernel: measure the calls after it
                warm = buf.out()
                ctx.check(lib.iso_extract_mc_noalloc(ctx.handle, C.byref(d), C.byref(warm)),
                          "warm-up")
            ctx.synchronize()
            before = free_device_bytes()
            out = buf.out()
            ctx.check(lib.iso_extract_mc_noalloc(ctx.handle, C.byref(d), C.byref(out)),
                      "iso_extract_mc_noalloc")
            after = free_device_bytes()
            print(f"{shape}: V={nv} T={ni // 3} free before {before} after {after}", flush=True)
            assert (int(out.vertex_count), int(out.index_count)) == (nv, ni)
            v, i = buf.fetch(out)
            assert v.tobytes() == ref.vertices.tobytes()
            assert np.array_equal(i, ref.indices)
            assert after == before
            # too small: counts reported, ISO_ERR_CAPACITY
            small = A.IsoMeshOut(buf.v.value, buf.i.value, nv // 2, ni // 2, 0, 0,
                                 A.ISO_MEMORY_DEVICE, 0)
            assert lib.iso_extract_mc_noalloc(ctx.handle, C.byref(d), C.byref(small)) == \
                A.ISO_ERR_CAPACITY
            assert (int(small.vertex_count), int(small.index_count)) == (nv, ni)
        finally:
            buf.free()


def test_noalloc_refuses_what_it_cannot_hold(ctx):
    lib = ctx.lib
    g = blob_grid(9, 9, 9, 3)
    ctx.upload_grid(g)
    out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_DEVICE, 0)
    d = desc_of(9, 9, 9)
    # host outputs and program sources are refused
    host = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
    assert lib.iso_extract_mc_noalloc(ctx.handle, C.byref(d), C.byref(host)) == \
        A.ISO_ERR_INVALID_ARGUMENT
    dp = desc_of(9, 9, 9)
    dp.source = A.ISO_SOURCE_PROGRAM
    assert lib.iso_extract_mc_noalloc(ctx.handle, C.byref(dp), C.byref(out)) == \
        A.ISO_ERR_INVALID_ARGUMENT
    # a lattice beyond the scratch: refused on the host, before any launch
    # (the binding only records dimensions; nothing reads the memory)
    dev = C.c_void_p()
    ctx.check(lib.iso_device_alloc(ctx.handle, 4 * 9 * 9 * 9, C.byref(dev)), "alloc")
    try:
        ctx.check(lib.iso_grid_bind_device(ctx.handle, dev, 1025, 1025, 1026), "bind")
        big = desc_of(1025, 1025, 1026)
        assert lib.iso_extract_mc_noalloc(ctx.handle, C.byref(big), C.byref(out)) == \
            A.ISO_ERR_INVALID_ARGUMENT
        assert b"no-allocation scratch" in lib.iso_ctx_last_error(ctx.handle)
    finally:
        ctx.upload_grid(g)
        lib.iso_device_free(ctx.handle, dev)
