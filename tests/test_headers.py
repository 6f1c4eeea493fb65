"""Every public header of the C++ drop-in compiles on its own (a translation
unit that includes only it), as the reference's callers include them."""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HEADERS = sorted(os.path.basename(p) for p in glob.glob(os.path.join(ROOT, "include", "*"))
                 if p.endswith((".h", ".hpp", ".cuh")))


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("header", HEADERS)
def test_header_is_self_contained(header, tmp_path):
    src = tmp_path / "tu.hip"
    src.write_text('#include "%s"\nint main() { return 0; }\n' % header)
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-std=c++20", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
