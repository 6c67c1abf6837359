"""What the trained warm-start network gains over the method without it.

Sparse-view workload (256^3 @1 mm from 90 noisy views of 512^2), held-out phantom
seed, CG with the matched Joseph pair to a relative residual of 1e-3 from zero,
from FDK and from CNN(FDK) with the committed weights.  The learned start must
reach the tolerance in at most 0.65x the iterations of plain CG from zero (measured
0.49x on this seed) and in less wall time (its own FDK, CNN and A X0 cost included),
in fewer iterations than the FDK start, from a smaller residual than FDK's, and end
at a volume at least as close to the zero-start end point, and closer to the phantom,
than the FDK start's."""
import pytest

pytestmark = pytest.mark.gpu


def test_learned_warm_start_beats_plain_cg(hip):
    from rtk_amd import warmstart_train as W
    from rtk_amd.warmstart import WarmStartCNN, default_weights_path
    net = WarmStartCNN.load(default_weights_path())
    g = W.geometry()
    W.evaluate_seed(net, g, W.EVAL_SEED0 + 1)  # warm-up: kernels, caches, allocations
    r = W.evaluate_seed(net, g, W.EVAL_SEED0)
    print({k: (r[k]["iterations"], round(r[k]["seconds"], 3), r[k]["start_rel_residual"]) for k in ("zero", "fdk", "cnn")})
    assert all(r[k]["reached"] for k in ("zero", "fdk", "cnn"))
    assert r["cnn"]["start_rel_residual"] < r["fdk"]["start_rel_residual"]
    assert r["cnn"]["iterations"] <= 0.65 * r["zero"]["iterations"]
    assert r["cnn"]["iterations"] < r["fdk"]["iterations"]
    assert r["cnn"]["seconds"] < r["zero"]["seconds"]
    assert r["cnn"]["rms_diff_vs_zero_start"] <= r["fdk"]["rms_diff_vs_zero_start"]
    assert r["cnn"]["rmse_vs_phantom"] < r["fdk"]["rmse_vs_phantom"]
