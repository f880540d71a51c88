"""The reference-side binding (include/ccr_hip_reference.hpp): ccr::ExposureEngine, ccr::NettingSet
and ccr::ScenarioGenerator mapped onto the C structs of ccr_hip.h, returning a ccr::ExposureResult.

Needs the reference tree (its headers and sources are compiled by `make -C oracle binding`); skipped
where it is absent. CPU only: the binding's C entry is routed to the FP64 oracle, which takes the
same structs, and its profile, reporting indices and scenario grid must equal the reference's own
ExposureEngine::run bit for bit (src/exposure/exposure_engine.cpp:50-81)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
needs_reference = pytest.mark.skipif(not (REF / "src").is_dir(), reason="reference tree not present")


@needs_reference
def test_binding_hands_the_c_entry_the_reference_inputs():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "oracle", "binding"], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "binding_check")], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "binding: all cases bit-identical" in out.stdout
    assert out.stdout.count("bit-identical (max |dEE| 0.000e+00)") == 5
    assert "unsupported trade: std::invalid_argument" in out.stdout


@needs_reference
def test_binding_gpu_entry_compiles(tmp_path):
    """The GPU form, ccr::hip::run(ctx, engine, nettingSet, q, precision), against libccr_hip's C ABI."""
    src = tmp_path / "use.cpp"
    src.write_text('#include "ccr_hip.hpp"\n#include "ccr_hip_reference.hpp"\n'
                   "ccr::ExposureResult f(const ccr::ExposureEngine& e, const ccr::NettingSet& n) {\n"
                   "  ccr::hip::Context ctx(0);\n"
                   "  return ccr::hip::run(ctx.get(), e, n, 0.975, CCR_FP32);\n}\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", f"-I{REF / 'include'}",
                    f"-I{ROOT / 'include'}", str(src)], check=True)
