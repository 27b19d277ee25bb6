"""HIP path against the oracle, through the C ABI (needs an MI355X).

Tolerances: seismograms rel-L2 <= 1e-5 and gradients
rel-L2 <= 1e-4 against the FP32 oracle; both also reported against FP64.
The per-step reference-structured kernels are compared element by element.
"""
import numpy as np
import pytest

from oracle import Oracle
from fwihip.workloads import Geometry, Shot, ricker, config1, layered_faulted, surface_shots

pytestmark = pytest.mark.gpu

SYN_TOL = 1e-5
GRAD_TOL = 1e-4


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from fwihip import lib as L
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


@pytest.fixture(scope="module")
def o32():
    return Oracle(np.float32)


@pytest.fixture(scope="module")
def o64():
    return Oracle(np.float64)


def small_geom():
    return Geometry(nx=150, nz=72, npml=10, dx=10.0, dz=10.0, dt=1e-3)


def small_case(nshots=3, nt=300, seed=0):
    g = small_geom()
    rng = np.random.default_rng(seed)
    vel = (2200 + 300 * rng.standard_normal((g.nz, g.nx))).astype(np.float32)
    vel = np.clip(vel, 1500, 3200).astype(np.float32)
    shots = []
    for j in range(nshots):
        irx = np.arange(j % 2, g.nx, 2, dtype=np.int32)
        ntj = nt - 37 * j  # uneven nt inside one batch
        shots.append(Shot(nt=ntj, isx=20 + 45 * j, isz=2 + j, wav=ricker(ntj, g.dt, 20.0), irx=irx,
                          irz=np.full(len(irx), 2 + (j % 3), np.int32)))
    return g, vel, shots


# ---------------------------------------------------------------- per step
def test_cpml_coe_matches_oracle(lib, o32):
    for npml, dx, dt in ((20, 10.0, 1e-3), (12, 12.5, 1e-3), (7, 5.0, 4e-4)):
        a, b = lib.cpml_coe(npml, dx, dt)
        ao, bo = o32.cpml_coe(npml, dx, dt)
        np.testing.assert_array_equal(a, ao)
        np.testing.assert_array_equal(b, bo)


def test_op_pml_steps_match_oracle(lib, o32):
    g = Geometry(nx=70, nz=40, npml=9, dx=10.0, dz=8.0, dt=1e-3)
    pnx, pnz, npml = g.pnx, g.pnz, g.npml
    coe = o32.fdcoe(g.dx, g.dz)
    tab = o32.cpml_tables(npml, g.dx, g.dz, g.dt)
    st = o32.new_pml_state(pnx, pnz, npml)
    dev = {k: lib.DevArray(v.shape) for k, v in st.items()}
    pd = lib.DevArray((pnz, pnx))
    rng = np.random.default_rng(11)
    for step in range(6):
        p = rng.standard_normal((pnz, pnx)).astype(np.float32)
        o32.op_pml(p, st, coe, tab, npml)
        pd.upload(p)
        lib.op_pml(g, pd, dev["u"], dev["v"], dev["ux"], dev["vz"], dev["u0"], dev["u1"], dev["v1"],
                   dev["ux0"], dev["ux1"], dev["vz1"])
        for k in st:
            got = dev[k].numpy()
            np.testing.assert_allclose(got, st[k], rtol=2e-6, atol=1e-6 * max(1.0, np.abs(st[k]).max()),
                                       err_msg=f"step {step} array {k}")


def test_op_interior_time_integral_bdr(lib, o32):
    g = Geometry(nx=60, nz=44, npml=6, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(12)
    coe = o32.fdcoe(g.dx, g.dz)
    p = rng.standard_normal((g.nz, g.nx)).astype(np.float32)
    u, v, ux, vz = o32.op_interior(p, coe)
    du, dv, dux, dvz = (lib.DevArray(a.shape) for a in (u, v, ux, vz))
    lib.op_interior(g, lib.DevArray(p.shape, p), du, dv, dux, dvz)
    for a, b in ((du, u), (dv, v), (dux, ux), (dvz, vz)):
        np.testing.assert_allclose(a.numpy(), b, rtol=2e-6, atol=1e-6)
    pold = rng.standard_normal(p.shape).astype(np.float32)
    ekp = rng.uniform(0.01, 0.1, p.shape).astype(np.float32)
    ref = o32.time_integral(p, pold, ux, vz, ekp)
    out = lib.DevArray(p.shape)
    args = [lib.DevArray(a.shape, a) for a in (p, pold, ux, vz, ekp)]
    lib.time_integral(out, *args)
    np.testing.assert_array_equal(out.numpy(), ref)
    pf = rng.standard_normal((g.pnz, g.pnx)).astype(np.float32)
    ud, lr = o32.write_bdr(pf, g.nx, g.nz, g.npml)
    dud = lib.DevArray((6, g.nx))
    dlr = lib.DevArray((g.nz, 6))
    lib.write_bdr(g, dud, dlr, lib.DevArray(pf.shape, pf))
    pb = lib.DevArray((g.nz, g.nx))
    lib.read_bdr(g, dud, dlr, pb)
    np.testing.assert_array_equal(dud.numpy(), ud)
    np.testing.assert_array_equal(dlr.numpy(), lr)
    np.testing.assert_array_equal(pb.numpy(), o32.read_bdr(ud, lr, np.zeros((g.nz, g.nx), np.float32)))


# ---------------------------------------------------------------- forward
def test_forward_batched_uneven(lib, o32, o64):
    g, vel, shots = small_case()
    ctx = lib.Context(g, max_shots=4, max_nt=300, max_nr=100, tile_z=8)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    syn = ctx.split(ctx.forward())
    for j, s in enumerate(shots):
        f32 = o32.forward(g, vel, s, want_strips=True)
        f64 = o64.forward(g, vel, s)
        print(f"shot {j}: syn rel32 {rel(syn[j], f32['syn']):.2e} rel64 {rel(syn[j], f64['syn']):.2e}")
        assert rel(syn[j], f32["syn"]) < SYN_TOL
        assert rel(syn[j], f64["syn"]) < SYN_TOL
        ps = ctx.read_state(4, j)
        xs = ctx.read_state(5, j)
        assert rel(ps, f32["pstrip"]) < SYN_TOL
        assert rel(xs, f32["xstrip"]) < SYN_TOL
        # final wavefields on the physical grid
        pl = f32["plast"]
        pn = ctx.read_state(s.nt & 1, j)[:g.nz, g.npml:g.npml + g.nx]
        assert rel(pn, pl[0]) < SYN_TOL


def test_forward_tiles_agree(lib):
    """Different tile heights, shot groups and the two forward kernel forms
    (z-marching, 2-D tile) give identical traces."""
    g, vel, shots = small_case(nshots=3, nt=200, seed=3)
    outs = []
    for tz, sg, kind in ((8, 1, 1), (13, 2, 1), (40, 4, 1), (0, 0, 1), (0, 0, 0)):
        ctx = lib.Context(g, max_shots=3, max_nt=200, max_nr=100, tile_z=tz, shots_per_group=sg)
        ctx.set_kernels(kind)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        outs.append(ctx.forward())
        ctx.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_tile_and_march_gradients_agree(lib):
    """The tile kernels (any reconstruction group size) and the z-marching
    kernels give the same traces bit for bit and the same gradient up to the
    FP64 order of the sum over shots."""
    g, vel, shots = small_case(nshots=3, nt=210, seed=4)
    rng = np.random.default_rng(6)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    outs = []
    for kind, rg in ((1, 0), (0, 1), (0, 2), (0, 8)):
        ctx = lib.Context(g, max_shots=3, max_nt=210, max_nr=100)
        ctx.set_kernels(kind)
        if rg:
            ctx.set_rec_group(rg)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        tr = ctx.forward()
        ctx.set_residual(ctx.pack(res))
        grad, gk = ctx.adjoint()
        outs.append((tr, grad, gk))
        ctx.close()
    for tr, grad, gk in outs[1:]:
        np.testing.assert_array_equal(tr, outs[0][0])
        print(f"tile vs march grad rel {rel(grad, outs[0][1]):.2e}")
        assert rel(grad, outs[0][1]) < 1e-12
        assert rel(gk, outs[0][2]) < 1e-12


def test_backward_chunks_give_the_same_image(lib):
    """The backward pass run in shot chunks (whole reconstruction groups per
    chunk, each chunk its own time loop over its own nt) gives the image of
    the single pass bit for bit: every group accumulator sees the same
    updates in the same order."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=13)
    rng = np.random.default_rng(14)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]

    def image(chunk, rg):
        ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
        ctx.set_rec_group(rg)
        ctx.set_bwd_chunk(chunk)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        ctx.forward(want_traces=False)
        ctx.set_residual(ctx.pack(res))
        gk = ctx.adjoint()[1]
        ctx.close()
        return gk

    for rg, chunks in ((1, (1, 2)), (2, (1, 2, 3))):
        ref = image(0, rg)
        for ch in chunks:   # rg 2: chunk 1 rounds up to 2 -> [0, 2), [2, 3); chunk 3 is one pass
            np.testing.assert_array_equal(image(ch, rg), ref)


@pytest.mark.parametrize("graph", [100, 7, 0])
def test_overlapped_backward_gives_the_same_image(lib, graph):
    """Adjoint step k beside imaging step k+1 on two streams (captured in
    graph blocks of any length, or plain launches) gives the image of the
    one-stream pass bit for bit, with uneven nt."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=15)
    rng = np.random.default_rng(16)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]

    def image(overlap):
        ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
        ctx.set_graph(graph)
        ctx.set_bwd_fused(False)
        ctx.set_bwd_overlap(overlap)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            ctx.forward(want_traces=False)
            ctx.set_residual(ctx.pack(res))
            out.append(ctx.adjoint()[1])
        ctx.close()
        return out

    ref = image(False)
    got = image(True)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ref[0], ref[1])


@pytest.mark.parametrize("graph", [100, 7, 0])
def test_fused_backward_gives_the_same_image(lib, graph):
    """One fused kernel per backward step (adjoint, then reconstruction and
    imaging from the adjoint's registers) gives the image of the separate
    kernels bit for bit: shot groups of 1, 2, 3 and 8 (a partial last
    group), uneven nt, backward chunks, an imaging window, graph replays."""
    g = Geometry(nx=190, nz=70, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(18)
    vel = (2200 + 300 * rng.standard_normal((g.nz, g.nx))).clip(1500, 3200).astype(np.float32)
    shots = []
    for j in range(5):
        irx = np.arange(j % 2, g.nx, 2, dtype=np.int32)
        ntj = 230 - 29 * j
        shots.append(Shot(nt=ntj, isx=15 + 40 * j, isz=2 + j, wav=ricker(ntj, g.dt, 20.0), irx=irx,
                          irz=np.full(len(irx), 2 + (j % 3), np.int32)))
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=5, max_nt=230, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def image(fused, rg, chunk=0, window=(1, 0, 2**31 - 1)):
        ctx.set_bwd_fused(fused)
        ctx.set_rec_group(rg)
        ctx.set_bwd_chunk(chunk)
        ctx.set_imaging(*window)
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            ctx.forward(want_traces=False)
            ctx.set_residual(ctx.pack(res))
            grad, gk = ctx.adjoint()
            out.append(gk)
        return out

    try:
        for rg in (1, 2, 3, 8):
            ref = image(False, rg)
            assert np.abs(ref[0]).max() > 0
            for a in ref[1:] + image(True, rg):
                np.testing.assert_array_equal(a, ref[0], err_msg=f"group {rg}")
        ref = image(False, 2, chunk=3, window=(3, 10, 200))
        for a in image(True, 2, chunk=3, window=(3, 10, 200)):
            np.testing.assert_array_equal(a, ref[0], err_msg="chunks + window")
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [100, 0])
def test_concurrent_general_strips_give_the_same_results(lib, graph):
    """General strips on a second stream beside the interior launch (forward
    and both backward forms, any cap on the interior workgroups per CU) give
    the one-stream traces and image bit for bit, over graph replays."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=19)
    rng = np.random.default_rng(20)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_rec_group(2)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run(concurrent, fused):
        ctx.set_concurrent(concurrent)
        ctx.set_bwd_fused(fused)
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            tr = ctx.forward()
            ctx.set_residual(ctx.pack(res))
            out.append((tr, ctx.adjoint()[1]))
        return out

    try:
        for fused in (True, False):
            ref = run(False, fused)
            for tr, gk in run(True, fused) + ref[1:]:
                np.testing.assert_array_equal(tr, ref[0][0])
                np.testing.assert_array_equal(gk, ref[0][1])
            for wgs in (0, 1, 2):               # interior occupancy caps (unused LDS reserved)
                ctx.set_interior_wgs(wgs)
                for tr, gk in run(True, fused):
                    np.testing.assert_array_equal(tr, ref[0][0])
                    np.testing.assert_array_equal(gk, ref[0][1])
            ctx.set_interior_wgs(3)
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [100, 0])
def test_strip_rows_give_the_same_results(lib, graph):
    """The tile kernels built with 4-row strips, in either pass or both, give
    the 8-row traces and image bit for bit: fused and separate backward steps,
    one and several streams, graph replays, and the checkpointed backward pass
    (its recompute runs the forward pass's strips).  The shallow sources and
    receivers put the free-surface rows, whose transposed rule reaches row 0
    from the second 4-row strip, into general strips of every kind."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=41)
    rng = np.random.default_rng(42)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_rec_group(2)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run():
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            tr = ctx.forward()
            ctx.set_residual(ctx.pack(res))
            out.append((tr, ctx.adjoint()[1]))
        return out

    try:
        for fused in (True, False):
            for concurrent in (True, False):
                ctx.set_bwd_fused(fused)
                ctx.set_concurrent(concurrent)
                ctx.set_pass_shape(0, 8, 3)
                ctx.set_pass_shape(1, 8, 3)
                ref = run()
                np.testing.assert_array_equal(ref[1][0], ref[0][0])
                np.testing.assert_array_equal(ref[1][1], ref[0][1])
                for rows in ((4, 4), (4, 8), (8, 4)):
                    ctx.set_pass_shape(0, rows[0], 5 if rows[0] == 4 else 3)
                    ctx.set_pass_shape(1, rows[1], 4 if rows[1] == 4 else 3)
                    for tr, gk in run():
                        np.testing.assert_array_equal(tr, ref[0][0], err_msg=f"{rows} fused={fused}")
                        np.testing.assert_array_equal(gk, ref[0][1], err_msg=f"{rows} fused={fused}")
        ctx.set_bwd_fused(True)
        ctx.set_concurrent(True)
        ctx.set_checkpointing(50)
        ctx.set_pass_shape(0, 8, 3)
        ctx.set_pass_shape(1, 8, 3)
        ref = run()
        for rows in ((4, 4), (4, 8), (8, 4)):
            ctx.set_pass_shape(0, rows[0], 3)
            ctx.set_pass_shape(1, rows[1], 3)
            for tr, gk in run():
                np.testing.assert_array_equal(tr, ref[0][0], err_msg=f"checkpointed {rows}")
                np.testing.assert_array_equal(gk, ref[0][1], err_msg=f"checkpointed {rows}")
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [100, 0])
def test_fwd_mixed_gives_the_same_results(lib, graph):
    """The forward step as one launch of interior and general strips gives the
    two-launch traces and image bit for bit, with either strip height, over
    graph replays, and in the checkpointed backward pass's recompute."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=45)
    rng = np.random.default_rng(46)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_rec_group(2)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run():
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            tr = ctx.forward()
            ctx.set_residual(ctx.pack(res))
            out.append((tr, ctx.adjoint()[1]))
        return out

    try:
        for ckpt in (0, 50):
            ctx.set_checkpointing(ckpt)
            for rows in (4, 8):
                ctx.set_pass_shape(0, rows, 3)
                ctx.set_fwd_mixed(False)
                ref = run()
                ctx.set_fwd_mixed(True)
                for tr, gk in run():
                    np.testing.assert_array_equal(tr, ref[0][0], err_msg=f"rows {rows} ckpt {ckpt}")
                    np.testing.assert_array_equal(gk, ref[0][1], err_msg=f"rows {rows} ckpt {ckpt}")
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [100, 7, 0])
def test_bwd_mixed_gives_the_same_image(lib, graph):
    """The fused backward step as one launch, whose general tiles image one
    step late, gives the image of the per-stream launches bit for bit: both
    strip heights, shot chunks (a chunk's image of step 0 inside or after its
    last graph block), uneven nt, an imaging window, over graph replays."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=47)
    rng = np.random.default_rng(48)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_rec_group(2)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run():
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            ctx.forward(want_traces=False)
            ctx.set_residual(ctx.pack(res))
            out.append(ctx.adjoint()[1])
        return out

    try:
        for rows in (4, 8):
            ctx.set_pass_shape(1, rows, 3)
            for chunk, window in ((0, (1, 0, 2**31 - 1)), (2, (1, 0, 2**31 - 1)), (0, (3, 5, 180))):
                ctx.set_bwd_chunk(chunk)
                ctx.set_imaging(*window)
                ctx.set_bwd_mixed(False)
                ref = run()
                ctx.set_bwd_mixed(True)
                for gk in run():
                    np.testing.assert_array_equal(gk, ref[0], err_msg=f"rows {rows} chunk {chunk} window {window}")
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [100, 50, 0])
def test_chunk_pairs_give_the_same_image(lib, graph):
    """Backward chunks run in pairs (the second chunk's steps on streams of
    its own beside the first's) give the image of chunks run one after
    another and of the single pass bit for bit: both strip heights, pairs
    whose chunks end at different steps (uneven nt), an imaging window, a
    last chunk without a partner, graph replays."""
    g = small_geom()
    rng = np.random.default_rng(61)
    vel = np.clip(2200 + 300 * rng.standard_normal((g.nz, g.nx)), 1500, 3200).astype(np.float32)
    shots = []
    for j in range(8):
        irx = np.arange(j % 2, g.nx, 2, dtype=np.int32)
        ntj = 230 - 17 * j
        shots.append(Shot(nt=ntj, isx=10 + 17 * j, isz=2 + j, wav=ricker(ntj, g.dt, 20.0), irx=irx,
                          irz=np.full(len(irx), 2 + (j % 3), np.int32)))
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=8, max_nt=230, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_rec_group(2)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run(n=8):
        out = []
        for _ in range(2):                      # a second pass replays the captured graphs
            ctx.forward(want_traces=False)
            ctx.set_residual(ctx.pack(res[:n]))
            out.append(ctx.adjoint()[1])
        return out

    try:
        for rows in (4, 8):
            ctx.set_pass_shape(1, rows, 3)
            ctx.set_bwd_chunk(0)
            ref = run()[0]
            assert np.abs(ref).max() > 0
            for chunk in (2, 4, 6):             # two pairs; one pair; a pair of 6 + 2 shots
                ctx.set_bwd_chunk(chunk)
                for lanes in (1, 2):
                    ctx.set_bwd_lanes(lanes)
                    for gk in run():
                        np.testing.assert_array_equal(gk, ref, err_msg=f"rows {rows} chunk {chunk} lanes {lanes}")
            ctx.set_bwd_lanes(2)
            ctx.set_imaging(3, 20, 190)          # an imaging window, pairs against one pass
            ctx.set_bwd_chunk(0)
            refw = run()[0]
            ctx.set_bwd_chunk(2)
            for gk in run():
                np.testing.assert_array_equal(gk, refw, err_msg=f"rows {rows} imaging window")
            ctx.set_imaging(1, 0, 2**31 - 1)
            ctx.set_shots(shots[:6])             # three chunks: a pair, then one alone
            ctx.set_bwd_chunk(0)
            ref6 = run(6)[0]
            ctx.set_bwd_chunk(2)
            for gk in run(6):
                np.testing.assert_array_equal(gk, ref6, err_msg=f"rows {rows} 6 shots chunk 2")
            ctx.set_shots(shots)
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [100, 0])
def test_tall_general_side_gives_the_same_image(lib, graph):
    """The fused backward step's general side on 8-row blocks (the 8-row
    kernels) beside 4-row interior tiles gives the image of the 4-row general
    side bit for bit: shots with a deep source and deep receiver rows (general
    blocks inside the grid, a block's 4-row tile beside an interior one), the
    grid's last partial block, chunks one at a time and in pairs, concurrent
    and one stream, and the setting refuses other heights."""
    g = small_geom()
    rng = np.random.default_rng(71)
    vel = np.clip(2200 + 300 * rng.standard_normal((g.nz, g.nx)), 1500, 3200).astype(np.float32)
    shots = []
    for j in range(6):
        irx = np.arange(j % 2, g.nx, 3, dtype=np.int32)
        ntj = 200 - 13 * j
        deep = j % 3 == 1
        shots.append(Shot(nt=ntj, isx=12 + 21 * j, isz=30 + j if deep else 2, wav=ricker(ntj, g.dt, 20.0),
                          irx=irx, irz=np.full(len(irx), 41 + j if deep else 3, np.int32)))
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=6, max_nt=200, max_nr=100)
    ctx.set_graph(graph)
    ctx.set_rec_group(2)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run():
        out = []
        for _ in range(2):
            ctx.forward(want_traces=False)
            ctx.set_residual(ctx.pack(res))
            out.append(ctx.adjoint()[1])
        return out

    try:
        for bad in (2, 6, 16):
            with pytest.raises(lib.FwiError):
                ctx.set_bwd_general_rows(bad)
        ctx.set_pass_shape(1, 4, 3)
        for concurrent in (True, False):
            ctx.set_concurrent(concurrent)
            for chunk, lanes in ((0, 2), (2, 2), (2, 1)):
                ctx.set_bwd_chunk(chunk)
                ctx.set_bwd_lanes(lanes)
                ctx.set_bwd_general_rows(4)
                ref = run()[0]
                assert np.abs(ref).max() > 0
                ctx.set_bwd_general_rows(8)
                for gk in run():
                    np.testing.assert_array_equal(
                        gk, ref, err_msg=f"concurrent {concurrent} chunk {chunk} lanes {lanes}")
    finally:
        ctx.close()


def test_setting_the_same_shots_again_keeps_the_results(lib):
    """set_shots with the batch the device already holds keeps its tables and
    captured graphs (a driver sets its shots before every gradient): traces
    and image are those of the first call bit for bit; a batch that differs
    only in a wavelet is taken up (doubled wavelet: doubled traces)."""
    g, vel, shots = small_case(nshots=3, nt=160, seed=81)
    rng = np.random.default_rng(82)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=160, max_nr=100)
    try:
        ctx.set_model(vel)

        def run(batch):
            ctx.set_shots(batch)
            tr = ctx.forward()
            ctx.set_residual(ctx.pack(res))
            return tr, ctx.adjoint()[1]

        tr0, gk0 = run(shots)
        for _ in range(2):
            tr, gk = run(shots)
            np.testing.assert_array_equal(tr, tr0)
            np.testing.assert_array_equal(gk, gk0)
        doubled = [Shot(nt=s.nt, isx=s.isx, isz=s.isz, wav=2 * s.wav, irx=s.irx, irz=s.irz) for s in shots]
        tr2, _ = run(doubled)
        assert rel(tr2, 2 * tr0) < 1e-6
        tr, gk = run(shots)
        np.testing.assert_array_equal(tr, tr0)
        np.testing.assert_array_equal(gk, gk0)
    finally:
        ctx.close()


def test_bwd_lanes_refuses_other_counts(lib):
    g, vel, shots = small_case(nshots=1, nt=50, seed=43)
    ctx = lib.Context(g, max_shots=1, max_nt=50, max_nr=100)
    try:
        for bad in (0, 3, -1):
            with pytest.raises(lib.FwiError):
                ctx.set_bwd_lanes(bad)
    finally:
        ctx.close()


def test_pass_shape_refuses_other_shapes(lib):
    g, vel, shots = small_case(nshots=1, nt=50, seed=43)
    ctx = lib.Context(g, max_shots=1, max_nt=50, max_nr=100)
    try:
        for args in ((0, 2, 3), (1, 16, 3), (0, 6, 3), (2, 8, 3), (0, 8, -1)):
            with pytest.raises(lib.FwiError):
                ctx.set_pass_shape(*args)
    finally:
        ctx.close()


@pytest.mark.parametrize("kind", [0, 1])
def test_imaging_window_is_a_partition_of_the_image(lib, kind):
    """The image over every step equals the sum of the images over the even
    and the odd steps (jump 2 from k0 = 0 and k0 = 1), and the sum of the
    images over [0, 99] and [100, end]; a window past the last step gives
    zero.  Reconstruction is unchanged, so traces and images add exactly up
    to the FP64 order of the sums."""
    g, vel, shots = small_case(nshots=2, nt=210, seed=11)
    rng = np.random.default_rng(12)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=2, max_nt=210, max_nr=100)
    ctx.set_kernels(kind)
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def image(jump=1, k0=0, k1=2**31 - 1):
        ctx.forward(want_traces=False)          # each adjoint consumes a forward
        ctx.set_residual(ctx.pack(res))
        ctx.set_imaging(jump, k0, k1)
        return ctx.adjoint()[1]

    full = image()
    even, odd = image(2, 0), image(2, 1)
    early, late = image(1, 0, 99), image(1, 100)
    none = image(1, 5000)
    with pytest.raises(lib.FwiError):
        ctx.adjoint()                           # no forward since the last adjoint
    ctx.close()
    assert np.abs(full).max() > 0 and not none.any()
    for a, b in ((even, odd), (early, late)):
        assert np.abs(a).max() > 0 and np.abs(b).max() > 0
        print(f"kind {kind}: split image rel {rel(a + b, full):.2e}")
        assert rel(a + b, full) < 1e-12


# ---------------------------------------------------------------- adjoint
def test_adjoint_matches_oracle(lib, o32, o64):
    g, vel, shots = small_case(nshots=2, nt=260, seed=5)
    ctx = lib.Context(g, max_shots=2, max_nt=260, max_nr=100, tile_z=16)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    ctx.forward(want_traces=False)
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx.set_residual(ctx.pack(res))
    grad, gk = ctx.adjoint()
    g32 = np.zeros_like(grad)
    g64 = np.zeros_like(grad)
    for s, r in zip(shots, res):
        f = o32.forward(g, vel, s, want_strips=True)
        g32 += o32.backward(g, vel, s, r, f)["grad"]
        f = o64.forward(g, vel, s, want_strips=True)
        g64 += o64.backward(g, vel, s, r.astype(np.float64), f)["grad"]
    print(f"adjoint grad rel32 {rel(grad, g32):.2e} rel64 {rel(grad, g64):.2e}")
    assert rel(grad, g32) < GRAD_TOL
    assert rel(grad, g64) < GRAD_TOL


def test_gradient_config1(lib, o32, o64):
    """Config 1: forward, L2 misfit on the device, adjoint, gradient."""
    g, vel, start, shots = config1()
    s = shots[0]
    obs = o64.forward(g, vel, s)["syn"].astype(np.float32)
    ctx = lib.Context(g, max_shots=1, max_nt=s.nt, max_nr=s.nr)
    ctx.set_model(start)
    ctx.set_shots([s])
    syn = ctx.forward().reshape(s.nr, s.nt)
    ctx.set_observed(obs.ravel())
    J = ctx.misfit_l2()
    grad, _ = ctx.adjoint()
    J32, syn32, gr32, _ = o32.gradient(g, start, s, obs)
    J64, syn64, gr64, _ = o64.gradient(g, start.astype(np.float64), s, obs.astype(np.float64))
    print(f"config1 syn rel32 {rel(syn, syn32):.2e} rel64 {rel(syn, syn64):.2e} "
          f"grad rel32 {rel(grad, gr32):.2e} rel64 {rel(grad, gr64):.2e} J {J} {J32} {J64}")
    assert rel(syn, syn32) < SYN_TOL and rel(syn, syn64) < SYN_TOL
    assert rel(grad, gr32) < GRAD_TOL and rel(grad, gr64) < GRAD_TOL
    assert abs(J - J64) < 1e-4 * J64


def test_batch_equals_sum_of_single_shots(lib):
    g, vel, shots = small_case(nshots=3, nt=220, seed=7)
    rng = np.random.default_rng(1)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=220, max_nr=100)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    ctx.forward(want_traces=False)
    ctx.set_residual(ctx.pack(res))
    gb, _ = ctx.adjoint()
    gs = np.zeros_like(gb)
    for s, r in zip(shots, res):
        ctx.set_shots([s])
        ctx.forward(want_traces=False)
        ctx.set_residual(r.ravel())
        gs += ctx.adjoint()[0]
    assert rel(gb, gs) < 1e-6


def test_reconstruction_recovers_initial_state(lib):
    """After the backward sweep the physical-grid field is p^0 = 0 (exact
    time reversal of the forward up to FP32 rounding)."""
    g, vel, shots = small_case(nshots=1, nt=300, seed=8)
    ctx = lib.Context(g, max_shots=1, max_nt=300, max_nr=100)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    ctx.forward(want_traces=False)
    peak = np.abs(ctx.read_state(300 & 1, 0)).max()
    ctx.set_residual(np.zeros(ctx.ntrace(), np.float32))
    ctx.adjoint()
    p0 = ctx.read_state(0, 0)[:g.nz, g.npml:g.npml + g.nx]
    assert np.abs(p0).max() < 1e-4 * peak


def test_graph_blocks_match_plain_launches(lib):
    """Captured blocks of steps (replayed many times, with eager work and
    synchronisation between the runs) give the traces and gradient of plain
    launches, bit for bit."""
    g, vel, shots = small_case(nshots=3, nt=230, seed=12)
    rng = np.random.default_rng(2)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    out = {}
    for steps in (0, 7, 100):
        ctx = lib.Context(g, max_shots=3, max_nt=230, max_nr=100)
        ctx.set_graph(steps)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        runs = []
        for rep in range(3):
            tr = ctx.forward()
            ctx.set_residual(ctx.pack(res))
            grad, _ = ctx.adjoint()
            ctx.synchronize()
            runs.append((tr, grad))
        for tr, grad in runs[1:]:
            np.testing.assert_array_equal(tr, runs[0][0])
            np.testing.assert_array_equal(grad, runs[0][1])
        out[steps] = runs[0]
        ctx.close()
    for steps in (7, 100):
        np.testing.assert_array_equal(out[steps][0], out[0][0])
        np.testing.assert_array_equal(out[steps][1], out[0][1])
