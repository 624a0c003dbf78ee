"""The C ABI library on the host (no GPU needed): it loads, exports every
symbol of include/necsim_hip.h, and its host-side scenario construction
(alias tables, origin sampling, seeding) matches the oracle bit for bit.
No device compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import necsim_hip as nh
import oracle as O
from necsim_hip import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "necsim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    syms = header_symbols()
    assert len(syms) >= 28
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/necsim_hip.h but not exported"
    assert set(syms) == set(_ffi.SIGNATURES), "ctypes signatures out of sync with the header"
    assert lib.nh_abi_version() == 1


def test_struct_layouts_match_reference_repr_c():
    assert nh.LINEAGE_DTYPE.itemsize == 32    # necsim/core/src/lineage.rs:97-109
    assert nh.EVENT_DTYPE.itemsize == 40      # necsim/core/src/event.rs:58-71 (size assert :68-71)
    assert nh.PACKED_EVENT_DTYPE.itemsize == 56  # necsim/core/src/event.rs:16-29 (size assert :38-44)
    assert _ffi.ATOM_DTYPE.itemsize == 24     # AliasMethodSamplerAtom<usize>
    assert _ffi.RANGE_DTYPE.itemsize == 16 and _ffi.SELF_DTYPE.itemsize == 16


def test_ctypes_structs_match_the_c_header(tmp_path):
    """sizes and field offsets of the ctypes mirrors against the C compiler's
    view of include/necsim_hip.h"""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    structs = {"nh_job_params": _ffi.JobParams, "nh_outcome": _ffi.Outcome, "nh_scenario": _ffi.Scenario,
               "nh_map_scenario": _ffi.MapScenario, "nh_server_params": _ffi.ServerParams,
               "nh_request": _ffi.Request, "nh_request_result": _ffi.RequestResult}
    dtypes = {"nh_lineage": nh.LINEAGE_DTYPE, "nh_speciation_event": nh.EVENT_DTYPE,
              "nh_packed_event": nh.PACKED_EVENT_DTYPE}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "necsim_hip.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    for name, dt in dtypes.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f in dt.names:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, f"{name}.{f}"
    for name, dt in dtypes.items():
        assert int(got[name]) == dt.itemsize, name
        for f in dt.names:
            assert int(got[f"{name}.{f}"]) == dt.fields[f][1], f"{name}.{f}"


def test_seed_from_u64_matches_oracle():
    for s in [0, 1, 7, 42, 2**63, 2**64 - 1]:
        assert _ffi.lib().nh_seed_from_u64(s) == O.lib().orc_seed_from_u64(s)


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("size,seed", [((8, 8), 5), ((40, 24), 3), ((128, 128), 42)])
def test_alias_tables_match_oracle(size, seed, packed):
    w_, h_ = size
    kw = dict(sigma=1.5, cutoff=2.5) if size == (8, 8) else {}
    hab, rp, col, wt = nh.generate_map(w_, h_, seed=seed, **kw)
    ranges, selfd, atoms = nh.build_alias_tables(w_, h_, hab, rp, col, wt, separable=not packed)
    t = O.Scenario.spatially_explicit(w_, h_, hab, rp, col, wt, packed=packed).tables()
    assert np.array_equal(ranges["start"], t["range_start"]) and np.array_equal(ranges["end"], t["range_end"])
    assert atoms["u"].tobytes() == t["atom_u"].tobytes()
    assert np.array_equal(atoms["e"], t["atom_e"]) and np.array_equal(atoms["k"], t["atom_k"])
    if not packed:
        assert selfd["self_dispersal"].tobytes() == t["self_p"].tobytes()
        assert np.array_equal(selfd["non_self_dispersal_event"], t["self_event"])


def test_alias_tables_contract_violation_is_an_error():
    hab = np.array([1, 0, 1, 1], np.uint32)
    rp = np.array([0, 2, 2, 3, 4], np.uint64)
    col = np.array([0, 1, 2, 3], np.uint32)  # cell 0 disperses to the uninhabitable cell 1
    w = np.ones(4)
    with pytest.raises(nh.NecsimHipError) as e:
        nh.build_alias_tables(2, 2, hab, rp, col, w)
    assert e.value.code == -5


@pytest.mark.parametrize("pct,rank,size", [(1.0, 0, 1), (0.37, 0, 1), (0.5, 2, 5), (0.01, 0, 1)])
def test_origin_sampling_matches_oracle(pct, rank, size):
    cases = [
        (nh.NonSpatialScenario((100, 100), 1).new(1e-3), O.Scenario.non_spatial(100, 100, 1), None),
        (nh.NonSpatialScenario((17, 9), 4).new(1e-3), O.Scenario.non_spatial(17, 9, 4), None),
        (nh.AlmostInfiniteNormalDispersalScenario(60, 10.0).new(1e-3), O.Scenario.almost_infinite_normal(10.0),
         (2**31 - 1, 2**31 - 1, 60)),
    ]
    hab, rp, col, wt = nh.generate_map(64, 48, seed=2)
    cases.append((nh.SpatiallyExplicitUniformTurnoverScenario(64, 48, hab, rp, col, wt).new(1e-3),
                  O.Scenario.spatially_explicit(64, 48, hab, rp, col, wt), None))
    for cogs, orc, circle in cases:
        got = cogs.sample_habitat(nh.OriginPreSampler.all().percentage(pct).partition(rank, size))
        want = orc.origins(pct, rank, size, circle=circle) if circle else orc.origins(pct, rank, size)
        assert got.tobytes() == want.tobytes()


def test_config_sizes():
    # lineage counts of the benchmark configurations
    assert len(nh.AlmostInfiniteNormalDispersalScenario(100, 10.0).new(1e-4)
               .sample_habitat(nh.OriginPreSampler.all())) == 31417
    hab, rp, col, wt = nh.generate_map(128, 128)
    assert int(hab.sum()) == 104294 and int((hab == 0).sum()) >= int(0.2 * 16384)


def test_map_generator_is_deterministic_golden():
    import hashlib
    hab, rp, col, wt = nh.generate_map(32, 32, seed=42)
    h = hashlib.sha256(hab.tobytes() + rp.tobytes() + col.tobytes() + wt.tobytes()).hexdigest()
    import json
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    assert h == golden["map32_sha256"]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_ffi, "_lib", None)
    monkeypatch.setattr(_ffi, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        _ffi.lib()
