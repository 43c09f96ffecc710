"""`portlbm --benchmark` (the reference's BENCHMARK_MODE driver,
src/lbm/benchmark/benchmark.cpp) on a small project: result files with the
reference's names and schema plus GPU time, MLUPS and GB/s per run, and the
phase-1 winners carried into phase 2."""
import json
import shutil
import subprocess
from pathlib import Path

import pytest

import portlbm_amd as pa

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


def test_three_phases_write_reference_schema(tmp_path):
    settings = tmp_path / "settings"
    settings.mkdir()
    doc = pa.settings("lptl", "stream", inner_h=128, inner_v=48, time_steps=60)
    (settings / "settings.json").write_text(json.dumps(doc))
    (settings / "benchmark.json").write_text(json.dumps({
        "testRunsPerConfig": 2, "testLinearTwoLattice": True, "dataLayouts": ["collision", "stream"],
        "algorithms": ["nptl", "nsol"], "phaseTwoScenarios": ["plate", "porous"]}))
    exe = ROOT / "portlbm_amd" / "bin" / "portlbm"
    out = subprocess.run([str(exe), "--benchmark", str(settings / "settings.json")], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert "Done with all benchmarks" in out.stdout
    wgs = [1024, 1023, 1022, 768, 767, 512, 511, 510, 256, 255, 254]
    for phase in (0, 1):
        files = sorted((tmp_path / "benchmarks" / f"phase{phase}").glob("*.json"))
        names = [f.name for f in files]
        assert len(files) == 3, names  # nptl, nsol, lptl
        for f in files:
            d = json.loads(f.read_text())
            assert f.name == f"results_phase{phase}_{d['algorithm']}_{d['device']}.json"
            assert (d["horizontalNodes"], d["verticalNodes"]) == (128, 48)  # as the settings file states them
            assert d["bytesPerFluidNodeUpdate"] == 144 and d["storage"] == "f64"
            for layout in ("collision", "stream"):
                entries = d[layout]
                if d["algorithm"] == "lptl":
                    assert [e["workGroupSize"] for e in entries] == ["auto"]
                else:
                    assert [e["workGroupSize"] for e in entries] == wgs
                for e in entries:
                    for k in ("initializationTimes", "runtimes", "kernelTimes", "mflups", "gbps"):
                        assert len(e[k]) == 2, (f.name, layout, k)
                    assert all(v > 0 for v in e["mflups"]) and all(v > 0 for v in e["kernelTimes"])
                    for m, g in zip(e["mflups"], e["gbps"]):
                        assert g == pytest.approx(m * 144 / 1000, rel=1e-9)
    p2 = sorted((tmp_path / "benchmarks" / "phase2").glob("*.json"))
    assert len(p2) == 6, [f.name for f in p2]  # (nptl, nsol, lptl) x (plate, porous)
    for f in p2:
        d = json.loads(f.read_text())
        assert f.name == f"phase2_{d['algorithm']}_{d['scenario']}_{d['device']}.json"
        assert d["dataLayout"] in ("collision", "stream")
        assert len(d["runtimes"]) == 2 and all(v > 0 for v in d["mflups"])
        if d["algorithm"] == "lptl":
            assert d["workGroupSize"] == "auto"
        else:
            assert d["workGroupSize"] in wgs
    assert "Optimal setup for algorithm nsol" in out.stdout
    # the settings file was rewritten for each configuration (as the reference does)
    last = json.loads((settings / "settings.json").read_text())
    assert last["domain"]["scenario"] == "porous"
