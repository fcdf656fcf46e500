"""GPU check of the C++ mirror include/drtopk.hpp: tests/cpp/hpp_mirror (built by
drtopksc_amd/build.py) calls every reference-named entry point (sample_radix_select,
radix_select, bucket_select, sample_bucket_select, bitonic_topk) as the reference's drivers
call theirs and compares each with drtopk_select on the same keys, bit for bit; with an empty
workspace each must refuse instead of allocating."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "hpp_mirror")


def test_hpp_mirror_equals_c_abi():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(EXE), "tests/cpp/hpp_mirror is not built: run python -m drtopksc_amd.build"
    print("[case] hpp mirror program", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
