"""The learned warm start at scale: 512^3 @0.5 mm from 180 noisy views of 1024^2
@0.4 mm (SID 1000 mm, SDD 1536 mm, 30 random ellipsoids, Poisson I0 = 5e3, held-out
seed).  The network runs as overlapping 128^3 tiles with 8-voxel halos; the test shows
that this equals whole-volume inference on the full 512^3 FDK volume (the 6-voxel
receptive radius lies inside the halo, and each output voxel's arithmetic does not
depend on the block it is computed in), and that CG (Joseph pair, FP32 projectors,
FP64 scalars) reaches a relative residual of 1e-3 from CNN(FDK) in fewer iterations
than from zero."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_at_scale_tiled_inference_and_warm_start(hip):
    from rtk_amd import warmstart_train as W
    from rtk_amd.warmstart import WarmStartCNN, scale_weights_path
    g, proj, vol, _ = W.scale_problem(W.EVAL_SEED0 + 100)
    x_fdk = W.fdk(g, proj, vol)
    net = WarmStartCNN.load(scale_weights_path())
    tiled = net.forward_tiled(x_fdk.tensor, tile=128, halo=8)
    whole = net(x_fdk.tensor)
    torch.cuda.synchronize()
    assert torch.isfinite(tiled).all()
    assert torch.equal(tiled, whole), float((tiled - whole).abs().max())
    del whole, net  # the whole-volume activations (43 GB) go with the network's workspace
    torch.cuda.empty_cache()
    _, r0, it0 = W.cg(g, proj, vol, 200, 1e-3)
    _, r1, it1 = W.cg(g, proj, vol.like(tiled), 200, 1e-3)
    print(f"at scale: CG to 1e-3 from zero {it0} iterations, from CNN(FDK) {it1} "
          f"(start residual {r1[0]:.3g})")
    assert r0[-1] <= 1e-3 and r1[-1] <= 1e-3
    assert it1 < it0
