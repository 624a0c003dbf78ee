"""Origin sampling on the device (nh_sample_origins_device, nh_job_add_origins)
against the oracle's origin samplers and the host walk (nh_sample_origins).

OriginPreSampler::all().percentage(p).partition(rank, size)
(origin_sampler/pre_sampler.rs:49-114) feeding NonSpatialOriginSampler
(non_spatial.rs:89-112), InMemoryOriginSampler (in_memory.rs:89-119) and
SingletonDemesCircleOriginSampler (singleton_demes/circle.rs:52-156).  Bar:
the lineage arrays are byte-identical to the oracle's, in order, for every
scenario kind, percentages, partitions and their combination; the
percentage pre-sampler's ranks are identical from the closed form of the
golden-ratio sequence and from its f64 loop; a simulation started from device origins
gives the oracle's events.  Device buffers come from the HIP runtime
(hipMalloc / hipMemcpy through ctypes).
"""
import ctypes

import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from necsim_hip import _ffi
from parity_utils import assert_exact, map_scenarios, oracle_run

pytestmark = pytest.mark.gpu

CENTRE = 2**32 // 2 - 1
D2H = 2


def announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


@pytest.fixture(scope="module")
def ctx():
    return nh.Context.get(0)


def device_origins(hip, ctx, tables, pre, circle=(0, 0, 0)):
    lib = _ffi.lib()
    n = ctypes.c_size_t(0)
    args = (ctx.handle, tables.handle, pre.proportion, pre.rank, pre.size, *circle)
    _ffi.check(lib.nh_sample_origins_device(*args, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, _ffi.LINEAGE_DTYPE)
    if n.value == 0:
        return out
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), out.nbytes) == 0
    try:
        _ffi.check(lib.nh_sample_origins_device(*args, d, n.value, ctypes.byref(n)))
        assert hip.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(d)
    return out


PRE_SAMPLERS = [nh.OriginPreSampler.all(), nh.OriginPreSampler.all().percentage(0.3),
                nh.OriginPreSampler.all().partition(2, 5), nh.OriginPreSampler.all().percentage(0.3).partition(3, 4),
                nh.OriginPreSampler.all().percentage(0.001), nh.OriginPreSampler.all().percentage(0.999),
                nh.OriginPreSampler.all().percentage(0.0), nh.OriginPreSampler.all().partition(7, 3)]


def pre_id(pre):
    return f"p{pre.proportion}-r{pre.rank}-s{pre.size}"


def check_same(hip, ctx, cogs, orc, pre, circle=None):
    tables = nh.Tables(ctx, cogs)
    circle_t = circle or (0, 0, 0)
    dev = device_origins(hip, ctx, tables, pre, circle_t)
    ref = orc.origins(pre.proportion, pre.rank, pre.size, circle=circle or (2**31 - 1, 2**31 - 1, 0))
    host = cogs.sample_habitat(pre)
    assert len(dev) == len(ref) == len(host), (len(dev), len(ref), len(host))
    assert dev.tobytes() == ref.tobytes(), "device origins differ from the oracle's"
    assert dev.tobytes() == host.tobytes()


@pytest.mark.parametrize("pre", PRE_SAMPLERS, ids=pre_id)
def test_non_spatial_origins(hip, ctx, pre):
    announce(f"device origins, non-spatial 100x100 deme 3, {pre_id(pre)}")
    cogs = nh.NonSpatialScenario((100, 100), 3).new(1e-3)
    check_same(hip, ctx, cogs, O.Scenario.non_spatial(100, 100, 3), pre)


@pytest.mark.parametrize("pre", PRE_SAMPLERS, ids=pre_id)
def test_circle_origins(hip, ctx, pre):
    announce(f"device origins, almost-infinite circle r=100, {pre_id(pre)}")
    scn = nh.AlmostInfiniteNormalDispersalScenario(100, 10.0)
    check_same(hip, ctx, scn.new(1e-4), O.Scenario.almost_infinite_normal(10.0), pre,
               circle=(CENTRE, CENTRE, 100))


@pytest.mark.parametrize("radius,centre", [(0, CENTRE), (1, CENTRE), (7, 3), (2000, CENTRE), (37, 2**32 - 5)])
def test_circle_shapes(hip, ctx, radius, centre):
    announce(f"device origins, circle radius {radius} at centre {centre} (wrapping extents included)")
    scn = nh.AlmostInfiniteNormalDispersalScenario(radius, 10.0, centre=(centre, centre))
    pre = nh.OriginPreSampler.all().percentage(0.01) if radius == 2000 else nh.OriginPreSampler.all()
    check_same(hip, ctx, scn.new(1e-4), O.Scenario.almost_infinite_normal(10.0), pre,
               circle=(centre, centre, radius))


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("pre", PRE_SAMPLERS, ids=pre_id)
def test_map_origins(hip, ctx, packed, pre):
    announce(f"device origins, 64x64 map packed={packed}, {pre_id(pre)}")
    h, orc = map_scenarios(64, 64, seed=42, packed=packed)
    check_same(hip, ctx, h.new(1e-3), orc, pre)


def test_config4_origins_full_size(hip, ctx):
    announce("device origins, config 4 circle r=2000 (12.6M lineages), all, against the host walk")
    scn = nh.AlmostInfiniteNormalDispersalScenario(2000, 50.0)
    cogs = scn.new(1e-6)
    tables = nh.Tables(ctx, cogs)
    dev = device_origins(hip, ctx, tables, nh.OriginPreSampler.all(), (CENTRE, CENTRE, 2000))
    host = cogs.sample_habitat(nh.OriginPreSampler.all())
    assert len(dev) == 12566345
    assert dev.tobytes() == host.tobytes()


@pytest.mark.parametrize("p", [1e-3, 0.1, 0.37, 0.5, 0.9, 0.999])
def test_presample_chunked_equals_serial_and_oracle(ctx, p):
    announce(f"percentage pre-sampler ranks p={p}: closed form vs the f64 loop vs the oracle")
    lib = _ffi.lib()
    n = 300_000
    chunked = np.zeros(n, np.uint64)
    serial = np.zeros(n, np.uint64)
    _ffi.check(lib.nh_debug_presample(ctx.handle, p, n, 0, _ffi.ptr(chunked)))
    _ffi.check(lib.nh_debug_presample(ctx.handle, p, n, 1, _ffi.ptr(serial)))
    assert chunked.tobytes() == serial.tobytes()
    assert np.all(np.diff(chunked.astype(np.int64)) >= 1)
    # the oracle's non-spatial origins with deme 1 carry the rank as their reference
    orc = O.Scenario.non_spatial(4000, 4000, 1).origins(p)
    m = min(n, len(orc))
    assert m > 1000
    assert np.array_equal(chunked[:m], orc["global_reference"][:m])


def test_presample_long_chain(ctx):
    announce("percentage pre-sampler: 3M ranks, closed form vs the f64 loop on one thread")
    lib = _ffi.lib()
    n = 3_000_000
    a = np.zeros(n, np.uint64)
    b = np.zeros(n, np.uint64)
    _ffi.check(lib.nh_debug_presample(ctx.handle, 0.75, n, 0, _ffi.ptr(a)))
    _ffi.check(lib.nh_debug_presample(ctx.handle, 0.75, n, 1, _ffi.ptr(b)))
    assert a.tobytes() == b.tobytes()


def job_for(ctx, tables, cogs, form=nh.PERSISTENT, pause=None):
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=form)
    return nh.Job(ctx, tables, args, nh.WyHash.seed_from_u64(42), cogs.speciation_probability, cogs.turnover_rate,
                  pause_before=pause)


@pytest.mark.parametrize("form", [nh.PERSISTENT, nh.PLAIN, nh.TASK_LIST])
def test_job_add_origins_lineages(ctx, form):
    announce(f"nh_job_add_origins, form {form}: the job's lineages before and after the start")
    cogs = nh.AlmostInfiniteNormalDispersalScenario(100, 10.0).new(1e-4)
    tables = nh.Tables(ctx, cogs)
    pre = nh.OriginPreSampler.all().percentage(0.5).partition(1, 3)
    ref = O.Scenario.almost_infinite_normal(10.0).origins(0.5, 1, 3, circle=cogs.circle)
    job = job_for(ctx, tables, cogs, form)
    try:
        job.add_origins(pre, cogs.circle)
        o = job.outcome()
        assert o.n_remaining == len(ref) and o.paused
        assert job.take_lineages().tobytes() == ref.tobytes()
        job.start()
        assert job.take_lineages().tobytes() == ref.tobytes()
    finally:
        job.close()


def test_initialise_from_device_origins_matches_oracle():
    announce("initialise_and_simulate (device origins), config 1 with percentage 0.5, against the oracle")
    cogs = nh.NonSpatialScenario((100, 100), 1).new(1e-3)
    orc = O.Scenario.non_spatial(100, 100, 1)
    lin = orc.origins(0.5)
    ref = oracle_run(orc, lin, 42, 1e-3, dedup="none")
    rep = nh.EventCollector()
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none())
    out = nh.HipIndependentAlgorithm.initialise_and_simulate(
        args, nh.WyHash.seed_from_u64(42), cogs, nh.OriginPreSampler.all().percentage(0.5), reporter=rep)
    assert_exact(ref, out, rep.events, "device origins")


def test_map_window_from_device_origins(ctx):
    announce("256x256 map, device origins at 20 %, window T=50, against a job fed by the host walk")
    h, _ = map_scenarios(256, 256, seed=42)
    cogs = h.new(1e-3)
    tables = nh.Tables(ctx, cogs)
    pre = nh.OriginPreSampler.all().percentage(0.2)
    streams = []
    for use_device in (True, False):
        job = job_for(ctx, tables, cogs, pause=50.0)
        try:
            if use_device:
                job.add_origins(pre)
            else:
                job.add_lineages(cogs.sample_habitat(pre))
            fin, ev = False, []
            while not fin:
                fin = job.step_round()
                ev.append(job.drain_events())
            o = job.outcome()
            streams.append((np.concatenate(ev).tobytes(), int(o.steps), o.time, job.take_lineages().tobytes()))
        finally:
            job.close()
    assert streams[0] == streams[1]


def test_add_origins_errors(ctx):
    announce("nh_job_add_origins: mixing with host lineages and oversized radii fail cleanly")
    cogs = nh.AlmostInfiniteNormalDispersalScenario(10, 10.0).new(1e-4)
    tables = nh.Tables(ctx, cogs)
    job = job_for(ctx, tables, cogs)
    try:
        job.add_origins(nh.OriginPreSampler.all(), cogs.circle)
        with pytest.raises(_ffi.NecsimHipError) as e:
            job.add_lineages(cogs.sample_habitat(nh.OriginPreSampler.all()))
        assert e.value.code == -4  # NH_E_STATE
        with pytest.raises(_ffi.NecsimHipError) as e:
            job.add_origins(nh.OriginPreSampler.all(), cogs.circle)
        assert e.value.code == -4  # NH_E_STATE
    finally:
        job.close()
    job = job_for(ctx, tables, cogs)
    try:
        with pytest.raises(_ffi.NecsimHipError) as e:
            job.add_origins(nh.OriginPreSampler.all(), (CENTRE, CENTRE, (1 << 22) + 1))
        assert e.value.code == -1  # NH_E_INVALID
    finally:
        job.close()
