"""The learned LW gas optics on the GPU: MFMA forward/backward against the
FP64 CPU reference (oracle/mlp_oracle.py), fused inference against the
training forward, training progress, and the table fallback outside the
training hull."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIZES = (7, 128, 128, 128, 256)


def _bf16(a):
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).bfloat16().float().numpy().astype(np.float64)


def _unflatten(flat):
    p, o = [], 0
    for k, n in zip(SIZES[:-1], SIZES[1:]):
        W = flat[o:o + k * n].reshape(k, n); o += k * n
        b = flat[o:o + n]; o += n
        p.append((W.astype(np.float64), b.astype(np.float64)))
    return p


def _emu(seed=3):
    from ccpprad import mlp
    e = mlp.LWEmulator(seed=seed)
    p, _ = e.get_params()
    rng = np.random.default_rng(seed)
    p = (p + rng.normal(0, 0.01, p.size)).astype(np.float32)   # non-zero biases
    e.set_params(p)
    return e, p


def test_mlp_forward_matches_fp64_reference():
    from oracle import mlp_oracle as M
    e, p = _emu()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(333, 7)).astype(np.float32)
    Y = torch.zeros((333, 256), dtype=torch.float32, device="cuda")
    e.forward(torch.from_numpy(X).cuda(), Y)
    torch.cuda.synchronize()
    params = [(_bf16(W), b) for W, b in _unflatten(p)]
    ref, _, _ = M.forward(params, X.astype(np.float64))
    err = np.abs(Y.cpu().numpy() - ref).max() / np.abs(ref).max()
    print("forward rel err", err)
    assert err < 2e-2


def test_mlp_gradients_match_fp64_reference():
    from oracle import mlp_oracle as M
    e, p = _emu()
    rng = np.random.default_rng(2)
    X = rng.normal(size=(1000, 7)).astype(np.float32)
    T = rng.normal(size=(1000, 256)).astype(np.float32)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    e.grads(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), loss)
    torch.cuda.synchronize()
    _, g = e.get_params()
    gg = _unflatten(g)
    params = [(_bf16(W), b) for W, b in _unflatten(p)]
    lref, gref = M.loss_and_grads(params, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss.item() - lref) < 1e-2 * lref
    for i in range(4):
        for j in range(2):
            a, b = gg[i][j].reshape(-1), gref[i][j].reshape(-1)
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            print(f"layer {i} {'Wb'[j]}: rel {rel:.2e} cos {cos:.6f}")
            assert rel < 3e-2 and cos > 0.999


def _profiles_dev(kd, ncol, nlay, seed):
    from ccpprad import synthetic as syn, capi, device
    pr = syn.make_profiles(ncol, nlay, seed=seed, clouds=False, kd=kd)
    arr = capi.profile_fields(pr)
    return pr, arr, device.to_device(arr)


def test_fused_inference_matches_training_forward(kd_lw):
    """The fused inference kernel and the training forward against the FP64
    forward of the same network (BF16 weights, unrounded activations).  Both
    GPU paths round every hidden activation to BF16 (8-bit significand), the
    fused kernel in its base-2 domain (softplus / ln2), so their roundings
    differ; each carries an error of a few BF16 ulp through the untrained
    network.  The fused path must be as accurate as the training path: its
    error on the standardised outputs within twice the training path's (plus
    1e-3 of the output scale), and below the absolute bounds."""
    from ccpprad import mlp
    from oracle import mlp_oracle as M
    e, p = _emu(5)
    ncol, nlay = 40, 127
    pr, arr, d = _profiles_dev(kd_lw, ncol, nlay, 9)
    F = mlp.features(pr)
    xm, xs = F.mean(0), F.std(0)
    rng = np.random.default_rng(0)
    ym, ys = rng.normal(-50, 1, 256), rng.uniform(0.5, 2, 256)
    ym[128:] = 0.1
    ys[128:] = 0.05
    e.set_norm(xm, xs, F.min(0) - 1, F.max(0) + 1, ym, ys)
    e.finalize()
    Xs = ((F - xm) / xs).astype(np.float32)
    X = torch.from_numpy(Xs).cuda()
    Y = torch.zeros((X.shape[0], 256), dtype=torch.float32, device="cuda")
    e.forward(X, Y)
    tau = torch.zeros((ncol, nlay, 128), dtype=torch.float32, device="cuda")
    pf = torch.zeros_like(tau)
    oob = torch.zeros((ncol * nlay,), dtype=torch.uint8, device="cuda")
    cf = torch.zeros((ncol,), dtype=torch.uint8, device="cuda")
    e.infer(ncol, nlay, d, tau, pf, oob, cf)
    torch.cuda.synchronize()
    params = [(_bf16(W), b) for W, b in _unflatten(p)]
    ref, _, _ = M.forward(params, Xs.astype(np.float64))         # standardised outputs, FP64
    cd = mlp.col_dry(pr.p_lev, pr.vmr["h2o"]).T.reshape(-1)
    y_inf = (np.log(tau.cpu().numpy().reshape(-1, 128) / cd[:, None]) - ym[:128]) / ys[:128]
    y_fwd = Y.cpu().numpy()[:, :128]
    scale = np.sqrt((ref[:, :128] ** 2).mean())
    e_inf = np.abs(y_inf - ref[:, :128]) / scale
    e_fwd = np.abs(y_fwd - ref[:, :128]) / scale
    pf_ref = np.maximum(ref[:, 128:] * ys[128:] + ym[128:], 0)
    d_inf = np.abs(pf.cpu().numpy().reshape(-1, 128) - pf_ref)
    d_fwd = np.abs(np.maximum(Y.cpu().numpy()[:, 128:] * ys[128:] + ym[128:], 0) - pf_ref)
    print("ln tau error / scale: fused max %.3g mean %.3g, training max %.3g mean %.3g" %
          (e_inf.max(), e_inf.mean(), e_fwd.max(), e_fwd.mean()))
    print("pfrac abs error: fused max %.3g, training max %.3g" % (d_inf.max(), d_fwd.max()))
    assert e_inf.mean() < 2 * e_fwd.mean() + 1e-3 and e_inf.max() < 2 * e_fwd.max() + 1e-3
    assert e_inf.max() < 5e-2 and e_inf.mean() < 5e-3
    assert d_inf.max() < 2 * d_fwd.max() + 1e-4 and d_inf.max() < 5e-3
    assert oob.sum().item() == 0


def test_training_reduces_loss_and_out_of_hull_falls_back(kd_lw, lut_lw):
    from ccpprad import mlp, lib, capi, device
    X, Y = mlp.make_training_set(kd_lw, nsamples=127 * 1600, torch=torch, chunk=1600)
    e = mlp.LWEmulator(seed=1)
    norm = mlp.fit_norm(X, Y)
    e.set_norm(**norm)
    hist = mlp.train(e, X, Y, torch, epochs=3, batch=4096, lr=2e-3)
    print("loss history", hist)
    assert hist[-1] < 0.5 * hist[0]
    # out-of-hull layer -> table values on that layer.  The hull is widened
    # to hold every layer of the test profiles, so that only the planted
    # layer lies outside it.
    ncol, nlay = 16, 127
    pr, arr, d = _profiles_dev(kd_lw, ncol, nlay, 33)
    F = mlp.features(pr)
    lo, hi = np.minimum(norm["lo"], F.min(0)), np.maximum(norm["hi"], F.max(0))
    norm["lo"] = (lo - 1e-3 * (hi - lo)).astype(np.float32)
    norm["hi"] = (hi + 1e-3 * (hi - lo)).astype(np.float32)
    e.set_norm(**norm)
    arr2 = dict(arr)
    arr2["t_lay"] = arr["t_lay"].copy()
    arr2["t_lay"][60, 3] = float(e.norm["hi"][0]) + 20.0      # hotter than anything seen in training
    d2 = device.to_device(arr2)
    s = lib.LWScheme(kd_lw, lut_lw, nlay)
    from oracle import oracle as O
    v = torch.from_numpy(O.vmr_stack(pr)).cuda()
    tau_t = torch.zeros((ncol, nlay, 128), dtype=torch.float32, device="cuda")
    pf_t = torch.zeros_like(tau_t)
    s.gas_optics_dev(ncol, nlay, d2["p_lay"], d2["p_lev"], d2["t_lay"], v, tau_t, pf_t)
    tau = torch.zeros_like(tau_t)
    pf = torch.zeros_like(tau_t)
    oob = torch.zeros((ncol * nlay,), dtype=torch.uint8, device="cuda")
    cf = torch.zeros((ncol,), dtype=torch.uint8, device="cuda")
    e.infer(ncol, nlay, d2, tau, pf, oob, cf)
    torch.cuda.synchronize()
    o = oob.cpu().numpy().reshape(ncol, nlay)
    assert o[3, 60] == 1 and o.sum() == 1 and cf.cpu().numpy()[3] == 1
    # through the scheme's gas-optics stage (the emulator attached): the
    # flagged layer carries exactly the table's optics, every other layer
    # exactly the emulator's
    e.attach(s)
    tau_s = torch.zeros_like(tau_t)
    pf_s = torch.zeros_like(tau_t)
    s.gas_optics_dev(ncol, nlay, d2["p_lay"], d2["p_lev"], d2["t_lay"], v, tau_s, pf_s)
    torch.cuda.synchronize()
    ts, ps = tau_s.cpu().numpy(), pf_s.cpu().numpy()
    tt, pt = tau_t.cpu().numpy(), pf_t.cpu().numpy()
    tn, pn = tau.cpu().numpy(), pf.cpu().numpy()
    assert np.array_equal(ts[3, 60], tt[3, 60]) and np.array_equal(ps[3, 60], pt[3, 60])
    keep = np.ones((ncol, nlay), dtype=bool)
    keep[3, 60] = False
    assert np.array_equal(ts[keep], tn[keep]) and np.array_equal(ps[keep], pn[keep])
    assert not np.array_equal(tn[3, 60], tt[3, 60])   # the emulator's own values there differ
    # and the scheme runs on that mix (a 3-epoch network: finite, positive
    # fluxes; the trained network's accuracy is test_gpu_mlp_gain's)
    ints = capi.lw_default_ints(ncol, nlay, nangle=3)
    full = device.to_device(arr2)
    full.update(device.lw_outputs(ncol, nlay))
    s.run_dev(ints, full)
    torch.cuda.synchronize()
    up_nn = full["fluxlwUP_allsky"].cpu().numpy()
    assert np.isfinite(up_nn).all() and (up_nn > 0).all()
    mlp.detach(s)
    tab = device.to_device(arr2)
    tab.update(device.lw_outputs(ncol, nlay))
    s.run_dev(ints, tab)
    torch.cuda.synchronize()
    d_up = np.abs(up_nn - tab["fluxlwUP_allsky"].cpu().numpy()).max()
    print("NN (3 epochs) vs table LW up flux max-abs", d_up)
