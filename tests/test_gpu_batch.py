"""Batched shots on the GPU.

* FP64 batch == single-shot FP64 path for every shot (<= 1e-10): the batch
  indexing (per-shot fields, materials, sources, seismograms) is exact.
* A shot's result does not depend on the batch it runs in (bitwise, FP32).
* FP32 batch vs FP64 per shot: relative L2 seismogram error <= 1e-4
  (the FP32 path's accuracy bound), reported per shot.
"""
import numpy as np
import pytest

from nexd_amd import batch, capi, layout

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return num / den


def fp64_single(shot_problem):
    arrs = layout.reference_arrays(shot_problem)
    with capi.Context(arrs, energy=False) as ctx:
        ctx.reset()
        ctx.run(1, shot_problem.nt)
        return ctx.seismograms()


def batch_run(sb, precision):
    with capi.BatchContext(sb, precision=precision) as b:
        b.reset()
        b.run(1, sb.base.nt)
        return [b.seismograms(k) for k in range(sb.nshot)]


@pytest.fixture(scope="module")
def cube_shots():
    return batch.multishot_cube(3, n=3, order=3, nt=300, full_source_arrays=True)


@pytest.fixture(scope="module")
def layered_shots():
    return batch.layered_shots(2, nx=4, ny=4, nz=3, order=4, nt=200, pml_cells=1, nrec_side=2,
                               full_source_arrays=True)


def test_fp64_batch_equals_single_shot(cube_shots):
    print(f"[gpu batch] cube shots={cube_shots.nshot} nelem={cube_shots.base.nelem}", flush=True)
    got = batch_run(cube_shots, capi.NEXD_FP64)
    for k, s in enumerate(cube_shots.shots):
        ref = fp64_single(s)
        assert np.max(rel_l2(got[k], ref)) <= 1e-10


def test_fp64_batch_equals_single_shot_attenuation_pml(layered_shots):
    print(f"[gpu batch] layered shots={layered_shots.nshot}", flush=True)
    got = batch_run(layered_shots, capi.NEXD_FP64)
    for k, s in enumerate(layered_shots.shots):
        ref = fp64_single(s)
        assert np.max(rel_l2(got[k], ref)) <= 1e-10


def test_shot_independent_of_batch(cube_shots):
    full = batch_run(cube_shots, capi.NEXD_FP32)
    one = batch.ShotBatch(cube_shots.base, [cube_shots.shots[1]], False)
    alone = batch_run(one, capi.NEXD_FP32)
    assert np.array_equal(full[1], alone[0])


def test_shot_independent_of_batch_pml(layered_shots):
    """The same with attenuation and PML (the FP32 contravariant volume
    term): a shot's seismograms do not depend on the other shots."""
    full = batch_run(layered_shots, capi.NEXD_FP32)
    one = batch.ShotBatch(layered_shots.base, [layered_shots.shots[1]], layered_shots.shared_material)
    alone = batch_run(one, capi.NEXD_FP32)
    assert np.array_equal(full[1], alone[0])


@pytest.mark.parametrize("which", ["cube", "layered"])
def test_fp32_batch_error_vs_fp64(which, cube_shots, layered_shots):
    sb = cube_shots if which == "cube" else layered_shots
    got = batch_run(sb, capi.NEXD_FP32)
    for k, s in enumerate(sb.shots):
        ref = fp64_single(s)
        # velocity and displacement traces carry the signal; accelerations
        # are checked too
        err = rel_l2(got[k], ref)
        print(f"  {which} shot {k}: max rel L2 FP32 vs FP64 = {err.max():.3e} (v: {err[:, 3:6].max():.3e})",
              flush=True)
        assert err.max() <= 1e-4


@pytest.mark.parametrize("which", ["cube", "layered"])
def test_fp32_compensated_error_vs_fp64(which, cube_shots, layered_shots):
    sb = cube_shots if which == "cube" else layered_shots
    with capi.BatchContext(sb, precision=capi.NEXD_FP32, compensated=True) as b:
        b.reset()
        b.run(1, sb.base.nt)
        got = [b.seismograms(k) for k in range(sb.nshot)]
    plain = batch_run(sb, capi.NEXD_FP32)
    for k, s in enumerate(sb.shots):
        ref = fp64_single(s)
        err, err0 = rel_l2(got[k], ref), rel_l2(plain[k], ref)
        print(f"  {which} shot {k}: FP32 compensated {err.max():.3e}, plain FP32 {err0.max():.3e}", flush=True)
        assert err.max() <= 1e-4


@pytest.mark.parametrize("timeint", [1, 3])
def test_fp64_batch_equals_single_shot_other_schemes(timeint):
    """Euler (buffer copy after the stage) and RK3 (three buffers, float
    weights) through the batch path against the single-shot path."""
    nt = 40 if timeint == 1 else 150
    # (40 Euler steps at this CFL grow slowly; both paths do the same arithmetic)
    sb = batch.multishot_cube(2, n=3, order=3, nt=nt, timeint=timeint, full_source_arrays=True)
    print(f"[gpu batch] timeint={timeint} shots={sb.nshot}", flush=True)
    got = batch_run(sb, capi.NEXD_FP64)
    for k, s in enumerate(sb.shots):
        ref = fp64_single(s)
        assert np.max(rel_l2(got[k], ref)) <= 1e-10


def test_fp32_compensated_rk3():
    """The RK3 branch of the compensated stage sum (weights 3/4, 1/4 and the
    float-rounded 1/3, 2/3)."""
    sb = batch.multishot_cube(2, n=3, order=3, nt=150, timeint=3, full_source_arrays=True)
    with capi.BatchContext(sb, precision=capi.NEXD_FP32, compensated=True) as b:
        b.reset()
        b.run(1, sb.base.nt)
        got = [b.seismograms(k) for k in range(sb.nshot)]
    for k, s in enumerate(sb.shots):
        err = rel_l2(got[k], fp64_single(s))
        print(f"  rk3 shot {k}: FP32 compensated {err.max():.3e}", flush=True)
        assert err.max() <= 1e-4


@pytest.mark.parametrize("order,attenuation", [(1, True), (2, True), (3, True), (5, True), (3, False)])
def test_fp32_batch_other_orders(order, attenuation):
    """FP32 batches at other orders (other block shapes, tilings and, in the
    PML, the contravariant volume term's node tiles and k-step chunks)
    against FP64 single shots, with PML, with and without attenuation."""
    sb = batch.layered_shots(2, nx=3, ny=3, nz=3, order=order, nt=80, pml_cells=1, nrec_side=2,
                             attenuation=attenuation, full_source_arrays=True)
    got = batch_run(sb, capi.NEXD_FP32)
    for k, s in enumerate(sb.shots):
        err = rel_l2(got[k], fp64_single(s))
        print(f"  N={order} att={attenuation} shot {k}: FP32 {err.max():.3e}", flush=True)
        assert err.max() <= 1e-4


def test_batch_energy_and_per_shot_pml_switch(layered_shots):
    """Each shot's energy after every step (src/timeloopMod.f90:508-516)
    and its own PML switch (src/timeloopMod.f90:532-539): FP64 batch with
    shot 1 switched off after step k, shot 0 left on, against the oracle
    run of each shot (seismograms and energy <= 1e-10); the FP32 batch's
    energies within 1e-4 of FP64."""
    from oracle import oracle
    sb = layered_shots
    nt, k = sb.base.nt, 80
    with capi.BatchContext(sb, precision=capi.NEXD_FP64, energy=True) as b:
        b.reset()
        b.run(1, k)
        b.set_pml_active(1, False)
        b.run(k + 1, nt)
        got = [(b.seismograms(i), b.energy(i)) for i in range(sb.nshot)]
    for i, s in enumerate(sb.shots):
        so, eo, _ = oracle.run(layout.reference_arrays(s), compat=3, pml_off_after=k if i == 1 else 0)
        assert np.max(rel_l2(got[i][0], so)) <= 1e-10
        assert np.all(eo > 0) and float(np.max(np.abs(got[i][1] - eo) / eo)) <= 1e-10
    with capi.BatchContext(sb, precision=capi.NEXD_FP32, energy=True) as b:
        b.reset()
        b.run(1, k)
        b.set_pml_active(1, False)
        b.run(k + 1, nt)
        e32 = [b.energy(i) for i in range(sb.nshot)]
    for i in range(sb.nshot):
        err = float(np.max(np.abs(e32[i] - got[i][1]) / got[i][1]))
        print(f"  shot {i}: FP32 energy vs FP64 max rel {err:.2e}", flush=True)
        assert err <= 1e-4


def test_many_receivers_span_several_step_end_blocks():
    """121 receivers, so one shot's 1,089 receiver values take two blocks
    of the step-end launch: the FP64 batch (2 shots x 2 blocks, the step
    counter advanced by the last block) and each shot's single-shot run (1 x
    2 blocks) against the oracle, seismograms and energy <= 1e-10."""
    from oracle import oracle
    sb = batch.layered_shots(2, nx=4, ny=4, nz=3, order=3, nt=120, pml_cells=1, nrec_side=11,
                             full_source_arrays=True)
    nrec = sb.base.rec_xyz.shape[0]
    print(f"[gpu batch] {nrec} receivers, {sb.nshot} shots", flush=True)
    assert nrec * 9 > 1024
    with capi.BatchContext(sb, precision=capi.NEXD_FP64, energy=True) as b:
        b.reset()
        b.run(1, 50)
        b.run(51, sb.base.nt)
        got = [(b.seismograms(i), b.energy(i)) for i in range(sb.nshot)]
    for i, s in enumerate(sb.shots):
        so, eo, _ = oracle.run(layout.reference_arrays(s), compat=3)
        assert np.all(eo > 0)
        assert np.max(rel_l2(got[i][0], so)) <= 1e-10
        assert float(np.max(np.abs(got[i][1] - eo) / eo)) <= 1e-10
        assert np.max(rel_l2(fp64_single(s), so)) <= 1e-10


def test_destroy_returns_device_memory():
    """A batch with PML that has run gives all its device memory back when
    destroyed: three create/run/destroy cycles lose less than 5 % of one
    batch's footprint (free memory from hipMemGetInfo of the HIP runtime the
    library runs on)."""
    import ctypes
    sb = batch.layered_shots(6, nx=10, ny=10, nz=5, order=4, nt=8, pml_cells=2, nrec_side=2, timeint=2,
                             attenuation=True)
    b = capi.BatchContext(sb, precision=capi.NEXD_FP32, energy=True)
    hip = ctypes.CDLL("libamdhip64.so.7")     # the runtime libnexd_gpu.so is linked against, already loaded

    def free():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    b.reset()
    b.run(1, 4)
    b.close()
    f0 = free()
    footprint = 0
    for _ in range(3):
        f1 = free()
        b = capi.BatchContext(sb, precision=capi.NEXD_FP32, energy=True)
        footprint = max(footprint, f1 - free())
        b.reset()
        b.run(1, 4)
        b.close()
    lost = f0 - free()
    print(f"[gpu batch] footprint {footprint / 2**20:.1f} MiB, lost after 3 cycles {lost / 2**20:.1f} MiB", flush=True)
    assert footprint > 0
    assert lost < 0.05 * footprint


def test_meshes_beyond_32_bit_planes_are_refused():
    """A mesh whose field plane exceeds 4 GiB is refused at creation with
    NEXD_E_UNSUPPORTED (the kernels address planes with 32-bit offsets),
    before anything is allocated."""
    import ctypes
    lib = capi._load_batch()
    nelem = 10_000_000     # N = 5: 5.6e8 nodes, 4.5 GB per FP64 plane
    cfg = capi.NexdConfig(nelem=nelem, order=5, nsrc=0, nrec=0, nt=10, timeint=2, attenuation=0, pml=0,
                          compat=capi.COMPAT_ALL, track_displacement=0, energy=0, device=0, dt=1e-3)
    h = ctypes.c_void_p()
    assert lib.nexd_create(ctypes.byref(cfg), ctypes.byref(h)) == -4
    assert b"4 GiB" in lib.nexd_last_error(None)
    bcfg = capi.NexdBatchConfig(nelem=nelem, order=5, nshot=1, nsrc=0, nrec=0, nt=10, timeint=2, attenuation=0,
                                pml=0, compat=capi.COMPAT_ALL, precision=capi.NEXD_FP32, shared_material=1,
                                compensated=0, energy=0, device=0, dt=1e-3)
    bcfg.nelem = 20_000_000    # FP32: 1.12e9 nodes, 4.5 GB per plane
    hb = ctypes.c_void_p()
    assert lib.nexd_batch_create(ctypes.byref(bcfg), ctypes.byref(hb)) == -4
    assert b"4 GiB" in lib.nexd_batch_last_error(None)
