"""Where the FP32- and FP64-forward gradients of the differentiation workload
part: the branches the two forward trajectories take differently.

The workload is the adjoint leg of bench.py: 96^3 dam break (liquid block
x<0.04, y<0.04, z<0.06 m), initial velocity = seeded Gaussian noise of std
1e-3 m/s smoothed over 4 cells, 400 steps, J = liquid-centroid height.  The
adjoint linearises every stage at the forward's states, so the FP32-forward
gradient dJ/du0 differs from the FP64-forward one wherever the two
trajectories take different branches: an upwind side, a limiter argument, a
clip.  The forward values of the two runs agree to O(1e-6); a flip replaces
the local Jacobian.

    python tests/branch_study.py gradients OUT.json   (GPU: the two gradients'
        relative L2 difference and the interior faces carrying most of it)
    python tests/branch_study.py perturbed OUT.json   (GPU: FP64 gradients
        whose initial velocity is rounded to FP32, or scaled by 1 + 1e-12
        noise, against the FP64 gradient: whether rounding-level changes of
        the input alone move the gradient as far as the FP32 forward does;
        the 8 seeded directional derivatives of each beside central
        differences of the FP64 forward)
    python tests/branch_study.py branches IN.json     (CPU oracle: both forward
        trajectories stage by stage, counting per step and per kind the
        faces or cells whose branch differs: overall, where the sweep's
        3-cell stencil holds both gas and liquid (F above 1e-3 and below
        1 - 1e-3), and within 3 cells of the faces IN.json lists)

Branch kinds (reference lines of voftrial1.cu):
  pred_upwind    predictor upwind side, sign of the advecting velocity of
                 each u*, v*, w* face along each axis (:266-275 and the v, w forms)
  fct_donor      sign of the face velocity in the donor-cell and downwind
                 fluxes, where F is nonzero on either side (:526-531)
  fct_clip1      stage-1 clip, Ftd outside [0, 1] before clamp01 (:532-535)
  fct_extrema    which of Ftd(c-1), Ftd(c), Ftd(c+1) is the max / the min in
                 the Zalesak bounds, ties a kind of their own (:551-552)
  fct_ratio      Zalesak ratio rp, rm: zero (no inflow), saturated at 1, or
                 the quotient (:570-582)
  fct_limiter    limiter coefficient: sign of the antidiffusive flux and which
                 ratio is the smaller (:600-604)
  fct_clamp4     stage-4 result outside [0, 1] before clamp01 (:640-645)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "cuda-lbm-ibm-fvm_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

N, STEPS, EVERY = 96, 400, 20
ORDER = ((0, 1, 2), (1, 2, 0), (2, 0, 1))  # sweep order by istep % 3 (:1384-1396)
KINDS = ("pred_upwind", "fct_donor", "fct_clip1", "fct_extrema", "fct_ratio", "fct_limiter", "fct_clamp4")


def initial_state(vof, va):
    """The adjoint leg's initial state (F of the box IC from the oracle, which
    the library's vof_init_box reproduces bit for bit)."""
    import oracle
    from vof_cases import SEED

    oracle.load()
    o = oracle.Oracle(vof.default_params(n=N, precision=8, g=(0.0, 0.0, -9.81), dt=4e-6))
    o.init_box(0.04, 0.04, 0.06)
    st = {"F": o.get("F"), "P": np.zeros(o.shape)}
    st.update({k: va.smoothed_noise(o.shape, 1e-3, 4.0, SEED + i) for i, k in enumerate("UVW")})
    o.close()
    return st


def gradients(out_path):
    import vof
    import vof_autograd as va

    vof.load_library()
    st = initial_state(vof, va)
    g = {}
    for prec in (4, 8):
        s = vof.Solver(vof.default_params(n=N, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6))
        for k, a in st.items():
            s.upload(k, a.astype(s.dtype))
        s.adj_forward(1, STEPS, EVERY)
        s.adj_backward(1.0)
        g[prec] = {k: s.adj_download(k) for k in "UVW"}
        s.close()
    d2 = {k: (g[4][k] - g[8][k]) ** 2 for k in "UVW"}
    tot = sum(float(a.sum()) for a in d2.values())
    den = sum(float((g[8][k] ** 2).sum()) for k in "UVW")
    flat = [(float(a[kk, jj, ii]), k, int(ii), int(jj), int(kk))
            for k, a in d2.items() for kk, jj, ii in zip(*np.unravel_index(np.argsort(a, axis=None)[-20:], a.shape))]
    flat.sort(reverse=True)
    top = [{"field": k, "i": i, "j": j, "k": kk, "share": v / tot} for v, k, i, j, kk in flat[:20]]
    res = {"rel_l2": float(np.sqrt(tot / den)), "top": top,
           "share_top9": sum(t["share"] for t in top[:9])}
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps({"rel_l2": res["rel_l2"], "share_top9": res["share_top9"]}), flush=True)


def perturbed(out_path):
    import vof
    import vof_autograd as va
    from vof_cases import SEED

    vof.load_library()
    st = initial_state(vof, va)

    def run(prec, state, every=EVERY, grad=True):
        s = vof.Solver(vof.default_params(n=N, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6))
        for k, a in state.items():
            s.upload(k, a.astype(s.dtype))
        J = s.adj_forward(1, STEPS, every)
        g = None
        if grad:
            s.adj_backward(1.0)
            g = {k: s.adj_download(k) for k in "UVW"}
        s.close()
        return J, g

    rng = np.random.default_rng(SEED + 77)
    cases = {
        "fp64": st,
        "fp64_u0_rounded_to_fp32": {**st, **{k: st[k].astype(np.float32).astype(np.float64) for k in "UVW"}},
        "fp64_u0_times_1+1e-12_noise": {**st, **{k: st[k] * (1 + 1e-12 * rng.standard_normal(st[k].shape))
                                                for k in "UVW"}},
        "fp32": st,
    }
    grads = {}
    for name, state in cases.items():
        print(f"{name}: forward + adjoint", flush=True)
        grads[name] = run(4 if name == "fp32" else 8, state)[1]
    g8 = grads["fp64"]
    den = sum(float((g8[k] ** 2).sum()) for k in "UVW")
    nu = float(np.sqrt(sum((st[k] ** 2).sum() for k in "UVW")))
    dirs = []
    for q in range(8):
        d = {k: va.smoothed_noise(st["U"].shape, 1.0, 4.0, SEED + 1000 + 3 * q + i) for i, k in enumerate("UVW")}
        nd = float(np.sqrt(sum((d[k] ** 2).sum() for k in "UVW")))
        dirs.append({k: a * (nu / nd) for k, a in d.items()})
    res = {"workload": f"{N}^3 dam break, u0 noise 1e-3 m/s, {STEPS} steps, checkpoint every {EVERY}",
           "gnorm_fp64": float(np.sqrt(den)), "cases": {}}
    for name, g in grads.items():
        d2 = {k: (g[k] - g8[k]) ** 2 for k in "UVW"}
        tot = sum(float(a.sum()) for a in d2.values())
        top = np.sort(np.concatenate([a.ravel() for a in d2.values()]))[::-1]
        res["cases"][name] = {
            "rel_l2_vs_fp64": float(np.sqrt(tot / den)),
            "share_of_gap_in_top9_faces": float(top[:9].sum() / tot) if tot > 0 else 0.0,
            "directional": [float(sum((g[k] * d[k]).sum() for k in "UVW")) for d in dirs]}
    h = 1e-4
    fd = []
    for q, d in enumerate(dirs):
        jp = run(8, {**st, **{k: st[k] + h * d[k] for k in "UVW"}}, STEPS, False)[0]
        jm = run(8, {**st, **{k: st[k] - h * d[k] for k in "UVW"}}, STEPS, False)[0]
        fd.append((jp - jm) / (2 * h))
        print(f"direction {q}: fd {fd[-1]:.10e}", flush=True)
    res["fd_fp64_h1e-4"] = fd
    for c in res["cases"].values():
        c["directional_rel_vs_fd"] = [abs(a - f) / abs(f) for a, f in zip(c["directional"], fd)]
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps({k: (c["rel_l2_vs_fp64"], max(c["directional_rel_vs_fd"]))
                      for k, c in res["cases"].items()}), flush=True)


def shift(a, axis, d):
    """b[c] = a[c + d e_axis] (edge cells keep their own value)."""
    b = np.roll(a, -d, axis=axis)
    if d > 0:
        idx = [slice(None)] * 3
        idx[axis] = slice(-d, None)
        b[tuple(idx)] = a[tuple(idx)]
    elif d < 0:
        idx = [slice(None)] * 3
        idx[axis] = slice(None, -d)
        b[tuple(idx)] = a[tuple(idx)]
    return b


def interior(shape):
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def pred_codes(U, V, W):
    """Sign (> 0) of the advecting velocity of every predictor face along each
    axis, in the reference's averages (axis order x, y, z; arrays [k, j, i])."""
    X, Y, Z = 2, 1, 0
    q = 0.25
    # u* face (i, j, k): ax = u, ay = avg v(i-1|i, j|j+1), az = avg w(i-1|i, k|k+1)
    uay = q * (shift(V, X, -1) + shift(shift(V, X, -1), Y, 1) + V + shift(V, Y, 1))
    uaz = q * (shift(W, X, -1) + shift(shift(W, X, -1), Z, 1) + W + shift(W, Z, 1))
    # v* face: ax = avg u(i|i+1, j-1|j), az = avg w(j-1|j, k|k+1)
    vax = q * (shift(U, Y, -1) + U + shift(shift(U, X, 1), Y, -1) + shift(U, X, 1))
    vaz = q * (shift(shift(W, Y, -1), Z, 1) + shift(W, Y, -1) + W + shift(W, Z, 1))
    # w* face: ax = avg u(i|i+1, k-1|k), ay = avg v(j|j+1, k-1|k)
    wax = q * (shift(shift(U, X, 1), Z, -1) + shift(U, Z, -1) + shift(U, X, 1) + U)
    way = q * (shift(shift(V, Y, 1), Z, -1) + shift(V, Z, -1) + V + shift(V, Y, 1))
    return [a > 0 for a in (U, uay, uaz, vax, V, vaz, wax, way, W)]


def sweep_codes(o, Fin, d):
    """Branch codes of the sweep just run along axis d (0 x, 1 y, 2 z)."""
    ax = 2 - d
    vel = o.get("UVW"[d])
    a = o.get("A" + "XYZ"[d])
    c = o.get("C" + "XYZ"[d])
    ftd = o.get("FTD")
    rp, rm = o.get("RP"), o.get("RM")
    s1 = o.get("S1")
    s4 = o.get("S4")
    active = (Fin != 0) | (shift(Fin, ax, -1) != 0)
    # interface: the 3-cell stencil along the sweep holds both gas and liquid
    fmx3 = np.maximum(Fin, np.maximum(shift(Fin, ax, -1), shift(Fin, ax, 1)))
    fmn3 = np.minimum(Fin, np.minimum(shift(Fin, ax, -1), shift(Fin, ax, 1)))
    codes = {"iface": ((fmx3 > 1e-3) & (fmn3 < 1 - 1e-3), None)}
    codes["fct_donor"] = (np.sign(vel).astype(np.int8), active)
    codes["fct_clip1"] = (((s1 > 1) | (s1 < 0)), None)
    fm, fp = shift(ftd, ax, -1), shift(ftd, ax, 1)
    mx = np.maximum(ftd, np.maximum(fm, fp))
    mn = np.minimum(ftd, np.minimum(fm, fp))
    ext = ((ftd == mx) * 1 + (fm == mx) * 2 + (fp == mx) * 4 + (ftd == mn) * 8 + (fm == mn) * 16 + (fp == mn) * 32)
    codes["fct_extrema"] = (ext.astype(np.int16), None)
    ratio = ((rp == 0) * 1 + (rp == 1) * 2 + (rm == 0) * 4 + (rm == 1) * 8)
    codes["fct_ratio"] = (ratio.astype(np.int8), None)
    # face f = low face of cell f: a >= 0 -> min(rp(f), rm(f-1)), else min(rp(f-1), rm(f))
    rpm, rmm = shift(rp, ax, -1), shift(rm, ax, -1)
    pos = a >= 0
    first = np.where(pos, rp, rpm)
    second = np.where(pos, rmm, rm)
    lim = pos * 1 + (first < second) * 2 + (first == second) * 4
    codes["fct_limiter"] = (lim.astype(np.int8), None)
    codes["fct_clamp4"] = (((s4 > 1) | (s4 < 0)), None)
    return codes


def run_stages(o, step, on_sweep=None):
    """One step of the oracle stage by stage (voftrial1.cu:1361-1401); the
    oracle's "jacobi" stage is one sweep, the step runs ten.  on_sweep(d, Fin)
    is called after each directional sweep with F before it."""
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w") + ("jacobi",) * 10 + ("update", "bc_vel"):
        o.stage(name)
    for d in ORDER[step % 3]:
        Fin = o.get("F") if on_sweep else None
        o.stage("fct_" + "xyz"[d])
        if on_sweep:
            on_sweep(d, Fin)
    o.stage("clamp")
    o.stage("bc_F")


def branches(in_path, steps=STEPS):
    import oracle
    import vof
    import vof_autograd as va

    top = json.load(open(in_path))["top"][:9] if in_path and os.path.exists(in_path) else []
    oracle.load()
    st = initial_state(vof, va)
    orcs = {}
    for prec in (4, 8):
        o = oracle.Oracle(vof.default_params(n=N, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6))
        for k, a in st.items():
            o.set(k, a)
        o.stats_enable(True)
        orcs[prec] = o
    # the stage sequence below is the oracle's step: check it on 3 steps
    for prec in (4, 8):
        pair = []
        for staged in (False, True):
            o = oracle.Oracle(vof.default_params(n=N, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6))
            for k, a in st.items():
                o.set(k, a)
            for step in range(1, 4):
                if staged:
                    run_stages(o, step)
                else:
                    o.step(step, 1)
            pair.append(o)
        assert all(np.array_equal(pair[0].get(f), pair[1].get(f)) for f in "FUVWP"), "staged step differs"
        for o in pair:
            o.close()
    shape = orcs[4].shape
    inner = interior(shape)
    near = np.zeros(shape, bool)
    for t in top:
        near[max(t["k"] - 3, 0):t["k"] + 4, max(t["j"] - 3, 0):t["j"] + 4, max(t["i"] - 3, 0):t["i"] + 4] = True
    total = {k: 0 for k in KINDS}
    at_iface = {k: 0 for k in KINDS}
    close = {k: 0 for k in KINDS}
    first_close = {k: None for k in KINDS}
    per_step = []

    def tally(step, kind, c4, c8, mask, iface=None):
        diff = (c4 != c8) & inner
        if mask is not None:
            diff &= mask
        n, m = int(diff.sum()), int((diff & near).sum())
        total[kind] += n
        if iface is not None:
            at_iface[kind] += int((diff & iface).sum())
        close[kind] += m
        if m and first_close[kind] is None:
            first_close[kind] = step
        return n

    for step in range(1, steps + 1):
        row = {k: 0 for k in KINDS}
        vel = {p: [o.get(f) for f in "UVW"] for p, o in orcs.items()}
        p4, p8 = pred_codes(*vel[4]), pred_codes(*vel[8])
        for a4, a8 in zip(p4, p8):
            row["pred_upwind"] += tally(step, "pred_upwind", a4, a8, None)
        cs = {4: [], 8: []}
        for p, o in orcs.items():
            run_stages(o, step, lambda d, Fin, p=p, o=o: cs[p].append(sweep_codes(o, Fin, d)))
        for q in range(3):
            for kind in KINDS[1:]:
                c4, m4 = cs[4][q][kind]
                c8, m8 = cs[8][q][kind]
                mask = None if m4 is None else (m4 | m8)
                row[kind] += tally(step, kind, c4, c8, mask, cs[4][q]["iface"][0] | cs[8][q]["iface"][0])
        per_step.append(row)
        if step % 50 == 0:
            print(json.dumps({"step": step, "flips_so_far": total}), flush=True)
    early = {k: sum(r[k] for r in per_step[:20]) for k in KINDS}
    res = {"workload": f"{N}^3 dam break, u0 noise 1e-3 m/s, {steps} steps (the adjoint leg)",
           "flips_total": total, "flips_at_interface": at_iface, "flips_first_20_steps": early,
           "flips_within_3_cells_of_top9_gap_faces": close, "first_step_near_top9": first_close,
           "top9": top}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "gradients":
        gradients(sys.argv[2])
    elif len(sys.argv) >= 3 and sys.argv[1] == "perturbed":
        perturbed(sys.argv[2])
    elif len(sys.argv) >= 2 and sys.argv[1] == "branches":
        branches(sys.argv[2] if len(sys.argv) > 2 else None, int(sys.argv[3]) if len(sys.argv) > 3 else STEPS)
    else:
        print(__doc__)
