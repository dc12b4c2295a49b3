"""A Matrix Market file through the HIP kernels (load -> SpMV / SpMM, non
and trans) against torch's float64 sparse product on the device."""

import numpy as np
import pytest
import torch

from flagsparse_amd import mtx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mtx_file(tmp_path_factory):
    rng = np.random.default_rng(3)
    n = 3000
    r = rng.integers(1, n + 1, 40000)
    c = rng.integers(1, n + 1, 40000)
    v = rng.standard_normal(40000)
    lines = ["%%MatrixMarket matrix coordinate real symmetric", f"{n} {n} {r.size}"]
    lines += [f"{max(a, b)} {min(a, b)} {float(x)!r}" for a, b, x in zip(r, c, v)]
    path = tmp_path_factory.mktemp("mtx") / "sym.mtx"
    path.write_text("\n".join(lines) + "\n")
    return str(path)


@pytest.mark.parametrize("kind", ["spmv", "spmm"])
@pytest.mark.parametrize("op", ["non", "trans"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_run_one_mtx(mtx_file, kind, op, dtype):
    res = mtx.run_one_mtx(mtx_file, dtype, torch.int32, warmup=1, iters=3, op=op, kind=kind, n_dense=32)
    assert res["status"] == "PASS", res
    assert res["nnz"] > 40000  # mirrored entries added


def test_cli(mtx_file, capsys):
    assert mtx.main([mtx_file, "--warmup", "1", "--iters", "2"]) == 0
    assert "PASS" in capsys.readouterr().out
