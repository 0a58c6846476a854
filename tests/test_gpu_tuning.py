"""The autotuner (pf_autotune): candidate launch choices timed on the plan's
own generation, the fastest kept, the on-disk cache hit on a second call, and
every choice it can make (tile width, pipeline chunks, generic tile lines and
radix orders, sub-batches) giving the oracle's fields."""
import numpy as np
import pytest

import oracle_py as orc
import pf_binding as pfb

pytestmark = pytest.mark.gpu


def desc(dim, cells, **kw):
    kw.setdefault("extensions", [1.0] * dim)
    return pfb.make_desc(dim, cells, **kw)


def rel_max(got, want):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(8)


def test_autotune_r2c_ensemble_and_cache(tmp_path):
    """A 128^3 Matern ensemble (FP32, 16 realizations): the tuner times the
    defaults and the tile-width / pipeline candidates, keeps a choice no
    slower than the defaults, the fields under it equal the default fields
    to FP32 rounding, and a second plan takes the choice from the cache."""
    import torch
    kw = dict(covariance="matern32", anisotropy="axiparallel", corr_length=[0.1, 0.1, 0.05],
              variance=2.0, approximate=True, precision="fp32")
    p = pfb.Plan(desc(3, [128, 128, 128], **kw))
    out = torch.empty((16, p.domain), dtype=torch.float32, device="cuda")
    ref = p.generate(16, seed=7, first=0, transform="logNormal",
                     out=torch.empty_like(out)).cpu().numpy()
    cache = str(tmp_path / "tuning.txt")
    r = p.autotune(16, out, seed=7, transform="logNormal", cache=cache)
    print("autotune:", r)
    assert not r["from_cache"] and r["candidates"] >= 5
    assert r["ms"] <= r["default_ms"]
    assert p.tuning() == r["best"]
    got = p.generate(16, seed=7, first=0, transform="logNormal", out=out).cpu().numpy()
    assert rel_max(np.log(got), np.log(ref)) < 1e-5
    p2 = pfb.Plan(desc(3, [128, 128, 128], **kw))
    r2 = p2.autotune(16, out, seed=7, transform="logNormal", cache=cache)
    assert r2["from_cache"] and r2["candidates"] == 0 and r2["best"] == r["best"]
    assert p2.tuning() == r["best"]
    p.close()
    p2.close()


@pytest.mark.parametrize("radix", [0, 1, 2])
@pytest.mark.parametrize("lines", [0, 2, 1])
def test_generic_kernels_under_every_radix_order_and_tile(radix, lines):
    """Lengths without specialised kernels (140 x 108 embedding: radices 2,
    5, 7 and 2, 3) run the generic kernels: each radix order and tile-line cap
    gives the oracle's FP64 field and product."""
    d = desc(2, [70, 54], covariance="exponential", corr_length=0.1)
    p = pfb.Plan(d)
    p.set_tuning(radix_order=radix, lines=lines)
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(3, npf)
    assert rel_max(p.generate(1, noise=noise)[0], orc.generate(d, noise)) < 1e-11
    x = np.random.default_rng(1).standard_normal(70 * 54)
    assert rel_max(p.matvec(x, "multiply"), orc.matvec(d, x)) < 1e-11
    p.close()


def test_autotune_generic_and_dctdst_and_dft(tmp_path):
    """Plans on the generic kernels, the DCT/DST field and the 1-D DFT field:
    the tuner explores tile lines / radix orders and sub-batches and the
    fields under the chosen configuration match the oracle."""
    import torch
    cases = [
        (2, [70, 54], dict(covariance="exponential", corr_length=0.1), 4),
        (3, [24, 20, 16], dict(covariance="exponential", corr_length=0.1, matrix="dct",
                               field="dctdst", approximate=True), 4),
        (1, [3000], dict(covariance="exponential", corr_length=0.05), 6),
    ]
    for dim, cells, kw, batch in cases:
        d = desc(dim, cells, **kw)
        p = pfb.Plan(d)
        out = torch.empty((batch, p.domain), dtype=torch.float64, device="cuda")
        r = p.autotune(batch, out, seed=5, cache=str(tmp_path / "t.txt"))
        print(dim, cells, r)
        assert r["candidates"] >= 3 and r["ms"] <= r["default_ms"]
        got = p.generate(batch, seed=5, first=0, out=out).cpu().numpy()
        npf = int(p.sizes.noise_per_field)
        if dim == 1:
            want = orc.generate(d, orc.noise_philox(5, 0, 0, npf))[0]
        else:
            want = orc.generate(d, orc.noise_philox(5, 0, 0, npf))
        assert rel_max(got[0], want) < 1e-11
        p.close()


def test_set_tuning_rejects_out_of_range():
    p = pfb.Plan(desc(2, [16, 16]))
    with pytest.raises(RuntimeError):
        p.set_tuning(radix_order=3)
    with pytest.raises(RuntimeError):
        p.set_tuning(chunks=9)
    p.close()
