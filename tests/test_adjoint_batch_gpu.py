"""The batched adjoint: 8 styles of the batch generator (128^3 float
densities) x 8 ring views at 512x512, 256 steps -- forward, L2 loss against
noisy-volume targets and adjoint of all 64 renders as one graph launch, the
gradients summed per style.  Images and records are bit-identical to the
per-render forward; gradients match the sum of per-render adjoints to float
summation order (relative L2 <= 1e-5); graph replays reproduce them."""
import pytest
import torch

from vkhr_amd.adjoint_batch import AdjointBatch, AdjointProblem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob(device):
    return AdjointProblem(device)


def _grads(prob):
    return [(st["d_density"].clone(), st["d_params"].clone()) for st in prob.styles]


def test_batched_adjoint_matches_per_render(prob):
    prob.run_plain()
    torch.cuda.synchronize()
    plain_imgs = [r["rgba"].clone() for r in prob.renders]
    plain_recs = [r["record"].clone() for r in prob.renders]
    plain = _grads(prob)
    batch = AdjointBatch(prob)
    batch.launch()
    torch.cuda.synchronize()
    for r, img, rec in zip(prob.renders, plain_imgs, plain_recs):
        assert torch.equal(r["rgba"], img)
        assert torch.equal(r["record"], rec)
        want = float(((img.double() - r["target"].double()) ** 2).sum())
        assert float(r["loss"]) == pytest.approx(want, rel=1e-4, abs=1e-6)
    for (gd, gp), (pd, pp) in zip(_grads(prob), plain):
        assert float(pd.abs().max()) > 0.0
        rel = float((gd.double() - pd.double()).norm() / pd.double().norm())
        assert rel <= 1e-5, rel
        assert torch.allclose(gp.double(), pp.double(), rtol=1e-4, atol=1e-6 * float(pp.abs().max()))
    first = _grads(prob)
    batch.launch()
    batch.launch()
    torch.cuda.synchronize()
    for (gd, gp), (fd, fp) in zip(_grads(prob), first):  # overwritten, not accumulated
        rel = float((gd.double() - fd.double()).norm() / fd.double().norm())
        assert rel <= 1e-5, rel


def test_batched_adjoint_time(prob):
    batch = AdjointBatch(prob)
    for _ in range(2):
        batch.launch()
        prob.run_plain()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    prob.run_plain()
    e[1].record()
    torch.cuda.synchronize()
    e[2].record(batch.stream)
    batch.launch(wait=False)
    e[3].record(batch.stream)
    torch.cuda.synchronize()
    plain_ms, batch_ms = e[0].elapsed_time(e[1]), e[2].elapsed_time(e[3])
    print(f"64 renders fwd+bwd: per-render {plain_ms:.2f} ms, one graph {batch_ms:.2f} ms "
          f"(x{plain_ms / batch_ms:.2f})")
    assert batch_ms < 1.5 * plain_ms
