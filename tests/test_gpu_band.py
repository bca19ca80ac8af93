"""The learned narrow band on the GPU: MFMA forward and gradients against
the CPU restatement (oracle/mlp.py), training, and band extraction checked
by the dense extraction (the mesh must be identical when no surface cell is
culled)."""
import numpy as np
import pytest

from isosurface_amd import MarchingCubes, _abi as A
from isosurface_amd import workloads as W
from isosurface_amd.band import (N_PARAMS, NarrowBand, brick_half_diagonal, brick_size,
                                  training_samples)
from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_parameter_roundtrip(ctx):
    band = NarrowBand(ctx)
    theta = M.init(11)
    band.set_params(theta)
    assert np.array_equal(band.get_params(), theta)
    band.init(3)
    assert len(band.get_params()) == N_PARAMS
    assert np.all(band.get_params()[:64 * 112].reshape(64, 112)[:, 99:] == 0.0)


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_restatement(ctx, seed):
    # Random weights; 4096 points; the device's MFMA products sum in another
    # order than the restatement, so a bf16 rounding of an activation can
    # flip: agreement to 1e-2 everywhere and 1e-4 at the median.
    band = NarrowBand(ctx)
    theta = M.init(seed)
    band.set_params(theta)
    p = np.random.default_rng(seed).random((4096, 3)).astype(np.float32)
    got = band.predict(p)
    want = M.forward(theta, p)
    err = np.abs(got - want)
    print("max", err.max(), "median", np.median(err), flush=True)
    assert np.max(err) <= 1e-2 * (1.0 + np.max(np.abs(want)))
    assert np.median(err) <= 1e-4 * (1.0 + np.median(np.abs(want)))


@pytest.mark.parametrize("overshoot", [1.0, 4.0])
def test_gradient_matches_restatement(ctx, overshoot):
    band = NarrowBand(ctx)
    theta = M.init(5)
    band.set_params(theta)
    rng = np.random.default_rng(5)
    n = 512
    p = rng.random((n, 3)).astype(np.float32)
    t = rng.normal(0, 0.05, n).astype(np.float32)
    w = (rng.random(n) + 0.5).astype(np.float32)
    loss, g = band.gradient(p, t, w, clamp=0.05, overshoot=overshoot)
    ref_loss, ref = M.loss_and_grad(theta, p, np.clip(t, -0.05, 0.05), w, overshoot)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    o = 0
    for name, shape in M.SIZES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], ref[o:o + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(name, rel, flush=True)
        assert rel <= 1e-2, name
        o += k
    # the padded feature columns stay exactly zero
    assert np.all(g[:64 * 112].reshape(64, 112)[:, 99:] == 0.0)


def test_training_fits_a_sphere(ctx):
    src = W.sphere_source(0.3)
    pts, sdf = training_samples(ctx, src, 60000, 60000, seed=1, coarse=65)
    band = NarrowBand(ctx)
    band.init(1)
    first = band.gradient(pts[:16384], sdf[:16384], clamp=0.05)[0]
    last = band.train(pts, sdf, steps=3000, batch=4096, lr=3e-3, clamp=0.05, seed=2)
    print("loss", first, "->", last, flush=True)
    assert last < 0.01 * first
    q = np.random.default_rng(3).random((20000, 3)).astype(np.float32)
    err = np.abs(band.predict(q) - np.clip(ctx.sample(src, q), -0.05, 0.05))
    print("median error", np.median(err), flush=True)
    # the same schedule in float64 on the CPU reaches about 7e-3 on this sphere
    assert np.median(err) < 1.5e-2


def band_setup(ctx, src, size, steps=3000):
    # A brick can only be culled when the clamped target can exceed its
    # half-diagonal: clamp at twice the half-diagonal.
    clamp = 2.0 * brick_half_diagonal(size)
    upper = (1.0, 1.0, size / (size - 1))
    pts, sdf = training_samples(ctx, src, 200000, 200000, seed=4, coarse=65, upper=upper)
    band = NarrowBand(ctx)
    band.init(4)
    w = (1.0 + 9.0 * np.exp(-np.abs(sdf) / 0.01)).astype(np.float32)
    band.train(pts, sdf, w, steps=steps, batch=8192, lr=3e-3, clamp=clamp, seed=5, overshoot=4.0)
    cal_p, cal_s = training_samples(ctx, src, 100000, 100000, seed=6, coarse=65, upper=upper)
    return band, band.calibrate(cal_p, cal_s, brick_half_diagonal(size))


@pytest.mark.parametrize("size", [65, 130])
def test_band_mesh_equals_dense(ctx, size):
    src = W.spheres_minus_torus()[0]
    band, margin = band_setup(ctx, src, size)
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    mesh, stats = band.extract_mesh(desc, margin)
    missed = band.missed_cells(desc)  # checks the last extraction: before the dense one
    dense = MarchingCubes(size, ctx).extract_mesh(src)
    print(f"size {size} margin {margin:.4f} band {stats.band_bricks}/{stats.bricks} "
          f"missed {missed}", flush=True)
    assert stats.band_bricks < stats.bricks
    assert missed == 0
    assert mesh.vertices.tobytes() == dense.vertices.tobytes()
    assert np.array_equal(mesh.indices, dense.indices)


@pytest.mark.parametrize("brick", [0, A.ISO_BAND_BRICK8])
def test_band_without_culling_is_dense(ctx, brick):
    src = W.sphere_source()
    band = NarrowBand(ctx)
    band.init(0)
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(40, A.ISO_SOURCE_PROGRAM)
    mesh, stats = band.extract_mesh(desc, 1e30, brick)
    per_axis = [-(-(n - 1) // brick_size(brick)) for n in (desc.nx, desc.ny, desc.nz)]
    assert stats.bricks == int(np.prod(per_axis))
    assert stats.band_bricks == stats.bricks
    assert band.missed_cells(desc) == 0
    dense = MarchingCubes(40, ctx).extract_mesh(src)
    assert mesh.vertices.tobytes() == dense.vertices.tobytes()
    assert np.array_equal(mesh.indices, dense.indices)
    # any other extraction replaces the sign words the dense check compares
    with pytest.raises(A.IsoError):
        band.missed_cells(desc)


def test_untrained_band_reports_its_misses(ctx):
    # An untrained network with no margin culls surface bricks: the missed
    # count is what the dense check finds, and only a mesh without misses
    # equals the dense one.
    src = W.sphere_source()
    band = NarrowBand(ctx)
    band.init(9)
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(48, A.ISO_SOURCE_PROGRAM)
    mesh, stats = band.extract_mesh(desc, 0.0)
    missed = band.missed_cells(desc)
    dense = MarchingCubes(48, ctx).extract_mesh(src)
    same = (len(mesh.indices) == len(dense.indices) and np.array_equal(mesh.indices, dense.indices)
            and mesh.vertices.tobytes() == dense.vertices.tobytes())
    # the dense check counts cells whose case differs: zero exactly when the meshes agree
    assert (missed == 0) == same


def certificate_centre_samples(program, size: int, brick: int, coarse: int = 4,
                               extra: float = 0.0) -> int:
    """Program evaluations of the two-level certificate over
    MarchingCubes::new(size), restated on the CPU: one per super-brick
    (coarse^3 bricks), plus one per brick of every super-brick whose centre
    value is within its half-diagonal (plus the fixed f32 allowances, plus
    `extra`) of zero.  The device adds bounds on the positions' and the
    evaluation's rounding, a few 1e-6 on these lattices, so its count lies
    between extra = 0 and extra = 2e-5."""
    d = O.reference_desc(size, 0)
    n = (d.nx, d.ny, d.nz)
    nb = [(k - 1 + brick - 1) // brick for k in n]
    nc = [(b + coarse - 1) // coarse for b in nb]
    span = brick * coarse
    f = np.float32
    count = 0
    for sz in range(nc[2]):
        for sy in range(nc[1]):
            for sx in range(nc[0]):
                c, p, d2 = (sx, sy, sz), [], f(0.0)
                for a in range(3):
                    lo = span * c[a]
                    hi = min(lo + span, n[a] - 1)
                    p.append(f(d.origin[a]) + f(0.5) * f(lo + hi) * f(d.spacing[a]))
                    e = f(span) * f(d.spacing[a])
                    d2 = f(d2 + e * e)
                count += 1
                hd = f(0.5) * np.sqrt(d2)
                if abs(O.sample(program, *p)) > hd * f(1.0 + 1e-4) + f(2e-5) + f(extra):
                    continue
                count += int(np.prod([min(coarse * c[a] + coarse, nb[a]) - coarse * c[a]
                                      for a in range(3)]))
    return count


@pytest.mark.parametrize("brick", [0, A.ISO_BAND_BRICK8])
@pytest.mark.parametrize("flags", [A.ISO_BAND_NO_NETWORK, A.ISO_BAND_CERTIFY])
def test_certified_band_is_dense(ctx, flags, brick):
    # The proofs alone (no network), or checking a network that predicts 1.0
    # everywhere, so that with no margin it culls every brick, surface ones
    # included: a brick leaves the band only by a proof, so none holds
    # surface, the mesh is the dense one, and far fewer samples than the
    # lattice's are evaluated.  Both give the same band.
    src = W.spheres_minus_torus()[0]
    band = NarrowBand(ctx)
    P = M.unpack(np.zeros(N_PARAMS, np.float32))
    P["b4"][:] = 1.0
    band.set_params(M.pack(P))
    ctx.upload_program(src.program())
    size = 130
    desc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    if flags == A.ISO_BAND_CERTIFY:
        _, unchecked = band.extract_mesh(desc, 0.0, brick)
        assert unchecked.band_bricks == 0
        assert band.missed_cells(desc) > 0  # what the proofs must catch
    mesh, stats = band.extract_mesh(desc, 0.0, flags | brick)
    missed = band.missed_cells(desc)
    dense = MarchingCubes(size, ctx).extract_mesh(src)
    print(f"band {stats.band_bricks}/{stats.bricks} exact {stats.exact_samples} "
          f"centres {stats.centre_samples} overturned {stats.overturned_culls} "
          f"missed {missed}", flush=True)
    lo = certificate_centre_samples(src.program(), size, brick_size(brick))
    hi = certificate_centre_samples(src.program(), size, brick_size(brick), extra=2e-5)
    assert lo <= stats.centre_samples <= hi
    assert stats.centre_samples < stats.bricks
    assert 0 < stats.band_bricks < stats.bricks // 2
    assert stats.exact_samples < 0.7 * size * size * (size + 1)
    if flags == A.ISO_BAND_CERTIFY:
        assert stats.overturned_culls == stats.band_bricks  # every band brick was a network cull
        _, plain = band.extract_mesh(desc, 0.0, A.ISO_BAND_NO_NETWORK | brick)
        assert (stats.band_bricks, stats.exact_samples) == (plain.band_bricks, plain.exact_samples)
    else:
        assert stats.overturned_culls == 0
    assert missed == 0
    assert mesh.vertices.tobytes() == dense.vertices.tobytes()
    assert np.array_equal(mesh.indices, dense.indices)


def test_certificate_refines_a_trained_band(ctx):
    src = W.spheres_minus_torus()[0]
    size = 130
    band, margin = band_setup(ctx, src, size)
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    _, plain = band.extract_mesh(desc, margin)
    mesh, cert = band.extract_mesh(desc, margin, A.ISO_BAND_CERTIFY)
    missed = band.missed_cells(desc)
    _, proofs = band.extract_mesh(desc, margin, A.ISO_BAND_NO_NETWORK)
    dense = MarchingCubes(size, ctx).extract_mesh(src)
    print(f"learned {plain.band_bricks}, certified {cert.band_bricks} of {cert.bricks}, "
          f"centres {cert.centre_samples}, overturned {cert.overturned_culls}", flush=True)
    # the proofs decide every brick: the band is the network-free one, and
    # the trained network's culls are, up to the few the proofs overturn, proven
    assert (cert.band_bricks, cert.centre_samples) == (proofs.band_bricks, proofs.centre_samples)
    # every band brick is one the network kept or a cull the proofs overturned
    assert cert.band_bricks <= plain.band_bricks + cert.overturned_culls
    assert cert.overturned_culls <= cert.band_bricks
    assert missed == 0
    assert mesh.vertices.tobytes() == dense.vertices.tobytes()
    assert np.array_equal(mesh.indices, dense.indices)


def test_band_error_paths(ctx):
    import ctypes as C
    lib = ctx.lib
    band = NarrowBand(ctx)
    band.init(0)
    desc = ctx.reference_desc(17, A.ISO_SOURCE_GRID)
    out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
    assert lib.iso_extract_mc_band(ctx.handle, C.byref(desc), 0.01, 0, C.byref(out), None) == \
        A.ISO_ERR_INVALID_ARGUMENT
    desc = ctx.reference_desc(17, A.ISO_SOURCE_PROGRAM)
    assert lib.iso_extract_mc_band(ctx.handle, C.byref(desc), -1.0, 0, C.byref(out), None) == \
        A.ISO_ERR_INVALID_ARGUMENT
    assert lib.iso_extract_mc_band(ctx.handle, C.byref(desc), 0.01, 8, C.byref(out), None) == \
        A.ISO_ERR_INVALID_ARGUMENT
    assert lib.iso_band_brick_size(0) == 4 and lib.iso_band_brick_size(A.ISO_BAND_BRICK8) == 8
    bad = np.zeros(10, np.float32)
    assert lib.iso_band_set_params(ctx.handle, bad.ctypes.data, 10) == A.ISO_ERR_INVALID_ARGUMENT
    p = np.zeros((8, 3), np.float32)
    t = np.zeros(8, np.float32)
    g = np.zeros(N_PARAMS, np.float32)
    assert lib.iso_band_gradient(ctx.handle, p.ctypes.data, t.ctypes.data, None, 8, 0.05, 1.0,
                                 g.ctypes.data, None) == A.ISO_ERR_INVALID_ARGUMENT
    p16 = np.zeros((16, 3), np.float32)
    t16 = np.zeros(16, np.float32)
    assert lib.iso_band_gradient(ctx.handle, p16.ctypes.data, t16.ctypes.data, None, 16, 0.05,
                                 0.5, g.ctypes.data, None) == A.ISO_ERR_INVALID_ARGUMENT


def test_certificate_at_a_far_origin(ctx):
    # A lattice at origin 4000.3 where a fixed f32 allowance proves bricks
    # that hold both signs (tests/test_band_certificate.py finds them on the
    # CPU): with the position and evaluation bounds no surface cell is lost
    # and the mesh is the dense one, for the proofs alone and for a network
    # that would cull every brick.
    from test_band_certificate import far_box
    prog, od = far_box()
    ctx.upload_program(prog)
    desc = A.IsoGridDesc()
    desc.nx, desc.ny, desc.nz = od.nx, od.ny, od.nz
    for i in range(3):
        desc.origin[i] = od.origin[i]
        desc.spacing[i] = od.spacing[i]
    desc.source = A.ISO_SOURCE_PROGRAM
    dense = ctx.extract_mc(desc)
    band = NarrowBand(ctx)
    P = M.unpack(np.zeros(N_PARAMS, np.float32))
    P["b4"][:] = 1.0
    band.set_params(M.pack(P))
    for flags in (A.ISO_BAND_NO_NETWORK, A.ISO_BAND_CERTIFY,
                  A.ISO_BAND_CERTIFY | A.ISO_BAND_BRICK8):
        mesh, stats = band.extract_mesh(desc, 0.0, flags)
        missed = band.missed_cells(desc)
        print(f"flags {flags}: band {stats.band_bricks}/{stats.bricks} missed {missed}", flush=True)
        assert missed == 0
        assert 0 < stats.band_bricks < stats.bricks
        assert mesh.vertices.tobytes() == dense.vertices.tobytes()
        assert np.array_equal(mesh.indices, dense.indices)
