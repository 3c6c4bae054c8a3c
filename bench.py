"""Benchmark of the CD3D hot path on one MI355X.

Workload (one "step"): 64^3 Q2 SUPG convection-diffusion (2,146,689 DOFs,
eps = 1e-4, b = curl of a seeded Fourier potential with max|b| = 1, f = 1,
homogeneous Dirichlet, 27-point rule), 5 GMG levels 4^3..64^3.  A step
assembles every level on the device and solves A x = b to FP64 relative
residual 1e-10 with V(2,2) multicolour-SSOR GMG preconditioning FGMRES(30):
  * fp64    : FP64 operator, MG and Krylov vectors (the plain method);
  * fp32    : FP32 operator/MG/Krylov vectors inside an FP64 defect
              correction (FP64 residual, Hessenberg/Givens, coarse solve);
  * learned : fp32 with learned smoothing on the levels below the finest
              (three network steps replace the pre- and the post-smoothing
              there; the finest level keeps SSOR).  The networks are trained
              once before the timed steps; training time is reported
              separately, not part of the step;
  * learned_fp16s : learned with the finest level's SSOR colour launches
              on row-scaled FP16 copies of the values (reported alongside;
              the metric's value is the learned form with FP32 storage).
After the timed forms the networks are trained again on other seeds of the
Fourier potential and the bench operator is solved with them (off-operator
networks), and the training-inclusive time of the learned form is reported.
The metric is the learned form's time to 1e-10 (seconds, lower is better),
with the plain FP64 time, assembly DOF/s and the SpMV roofline alongside.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nref 6] [--no-cpu]
    python bench.py --transient 200       # TCD3D 128^3 Q1 SUPG, Crank-Nicolson, 200 steps
    python bench.py --train-set 256 --train-steps 16   # networks across 256 seeded 64^3 Q1 fields

Hardware-counter passes (rocprofv3 --pmc) run tests/cpp/kernel_probe, a C++
program that makes the same launches on the same system.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector = MFMA), TFLOP/s
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA, TFLOP/s
# HBM bytes per dispatch at 64^3 Q2 from rocprofv3 FETCH_SIZE x 2.000 +
# WRITE_SIZE x 1.000 passes over tests/cpp/kernel_probe (the factors make the
# counters of a 1 GiB device copy read 1 GiB each), in
# profiles/k_rows_ssor_spmv_fp32_q2_64_hbm_traffic.md:
TRAFFIC_SSOR_LAUNCH = 30_917_171  # k_rows<float,3,1>, one colour launch (average over the sweeps)
TRAFFIC_SPMV = 561_289_188  # k_rows<float,0,1>


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(nref_cpu=5, seed=2024):
    """The oracle (single-threaded FP64 C++ restatement of the reference
    algorithms: linear-search CSR scatter, lexicographic SSOR, cellwise
    transfers, FGMRES with MGS, dense LU on the coarsest level) on the same
    workload at 32^3 Q2 (274,625 DOFs, 5 levels 2^3..32^3): set-up, assembly
    of every level and the FGMRES(30) V(2,2) SSOR solve to FP64 relative
    residual 1e-10, timed on one core of this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    from parmoon_amd import prm as P
    prm = P.config2(nref=nref_cpu, seed=seed)
    t0 = time.perf_counter()
    S = orc.System(nref_cpu - 4, 5, 2)
    t_setup = time.perf_counter() - t0
    fine = S.levels[-1]
    t0 = time.perf_counter()
    S.assemble(prm)
    t_asm_all = time.perf_counter() - t0
    rhs = S.rhs()
    sol0 = np.zeros(fine.N)
    sol0[fine.active:] = rhs[fine.active:]
    E = S.entries()
    r0 = np.linalg.norm(fine.defect(E, sol0, rhs)[0])
    opt = P.solver_opt(solver=16, restart=30, maxit=400, res_norm_min=1e-10 * r0, smoother=3, omega=1.0)
    t0 = time.perf_counter()
    s, info = S.solve(opt, rhs, sol0)
    t_solve = time.perf_counter() - t0
    return {"N": fine.N, "t_setup": t_setup, "t_assembly_all": t_asm_all, "t_solve": t_solve,
            "iters": int(info[0]), "final_rel_res": float(info[1] / r0)}


REF_SOLVE = os.path.join(ROOT, "oracle", "_ref", "ref_solve")


def cpu_baseline_reference(nref_cpu=5, seed=2024):
    """ParMooN's own CPU path (oracle/_ref/ref_solve: the reference's
    Assemble3D, TMultiGrid3D / TMGLevel3D with lexicographic SSOR,
    TMultiGridScaIte and TFgmresIte, compiled from its sources with g++ -O2)
    on the same workload at 32^3 Q2 (274,625 DOFs, 5 levels 2^3..32^3):
    set-up, assembly of every level and the FGMRES(30) V(2,2) SSOR solve to
    FP64 relative residual 1e-10, on one core of this host (the reference's
    sequential build).  The mesh is the reference's unit cube
    (tests/golden/Wuerfel.PRM / .GEO)."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_solve_fixtures as SF
    from parmoon_amd import prm as P
    prm = P.config2(nref=nref_cpu, seed=seed)
    opt = P.solver_opt(solver=16, restart=30, maxit=400, res_norm_min=0.0, red_factor=0.0, smoother=3, omega=1.0,
                       pre=2, post=2, rel_target=1e-10)
    mesh = (os.path.join(ROOT, "tests", "golden", "Wuerfel.PRM"), os.path.join(ROOT, "tests", "golden", "Wuerfel.GEO"))
    x, _, st = SF.run_reference(nref_cpu - 4, 5, 2, prm, opt, stats=True, exe=REF_SOLVE, mesh=mesh)
    return {"N": len(x), "t_setup": st["setup_s"], "t_assembly_all": st["assembly_s"], "t_solve": st["solve_s"],
            "iters": st["iterations"]}


def transient(args):
    """Transient workload: TCD3D on 128^3 Q1 (2,146,689 DOFs), eps = 1e-4,
    b = (-(y - 1/2), x - 1/2, 0.2), a seeded Gaussian blob as initial value,
    Crank-Nicolson with tau = 1e-3, SUPG (MatrixMRhsAssemble_SUPG at the
    start, MatricesAKRhsAssemble_SUPG every step, system matrix
    M + K + tau/2 A; --galerkin: the Galerkin forms); the stiffness matrices
    and rhs are re-assembled on the device every step and every step is solved
    to FP64 relative residual 1e-10 (5 GMG levels 8^3..128^3, V(2,2) SSOR
    with omega 0.4 for SUPG -- the time-derivative matrix K makes M + K far
    from diagonally dominant, and omega = 1 needs 3x the iterations -- or 1.0
    for Galerkin, FGMRES(50)).  Prints one JSON line with the per-form totals."""
    import parmoon_amd as pa
    from parmoon_amd import prm as P
    ctx = pa.Context(0)
    t0 = time.perf_counter()
    S = pa.System(ctx, 3, 5, 1)
    t_setup = time.perf_counter() - t0
    N = int(S.dims(-1)[0])
    supg = not args.galerkin
    prm = P.rotating(seed=1, form=P.FORM_TCD_SUPG_AK if supg else P.FORM_TCD_A)
    tau, steps = 1e-3, args.transient
    td = [tau, .5, .5, .5, .5]
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    prec_of = {"fp64": 0, "fp32": 1, "learned": 2, "learned_fp16s": 4}
    omega = args.omega if args.omega else (0.4 if supg else 1.0)
    base = dict(solver=16, restart=50, maxit=300, res_norm_min=0.0, red_factor=0.0, smoother=3, omega=omega,
                pre=2, post=2, rel_target=1e-10)
    res = {}
    train = None
    for form in forms:
        S.tcd_init(prm, None)
        if form in ("learned", "learned_fp16s") and train is None:
            # the system matrix of the first step, then the networks
            p1 = prm.copy()
            p1[P.TIME] = tau
            S.tcd_step(p1, td, P.solver_opt(**dict(base, precision=1)))
            tinfo = S.learned_train(P.solver_opt(**dict(base, precision=1)), samples=32, steps=args.train_steps,
                                    lr=2e-3, seed=1, levels=args.learned_levels, data=1, skip=args.learned_skip)
            train = {"seconds": round(float(tinfo[2]), 3), "loss_first": round(float(tinfo[0]), 4),
                     "loss_last": round(float(tinfo[1]), 4), **learned_config(args)}
            S.tcd_init(prm, None)
        opt = P.solver_opt(**dict(base, precision=prec_of[form], **learned_opts(args)))
        t, its, fb, t_solve = 0.0, 0, 0, 0.0
        t0 = time.perf_counter()
        for m in range(steps):
            t += tau
            p = prm.copy()
            p[P.TIME] = t
            info, rc = S.tcd_step(p, td, opt, check=False)
            if rc != 0:
                print(f"[bench] transient {form} step {m} rc={rc} info={info}", file=sys.stderr, flush=True)
                break
            its += int(info[0])
            fb += int(info[5])
            t_solve += info[4]
            if (m + 1) % 50 == 0:
                print(f"[bench] transient {form}: step {m + 1}/{steps}", file=sys.stderr, flush=True)
        ctx.sync()
        wall = time.perf_counter() - t0
        sol = S.solution()
        res[form] = {"total_s": round(wall, 4), "solve_s": round(t_solve, 4), "per_step_ms": round(wall / steps * 1e3, 3),
                     "iterations": its, "fallbacks": fb, "rc": int(rc), "final_max": float(np.max(sol))}
        print(f"[bench] transient {form}: {res[form]}", file=sys.stderr, flush=True)
    main_form = next(f for f in ("learned", "fp32", "fp64") if f in res)
    t64 = res.get("fp64", {}).get("total_s")
    out = {"metric": "TCD3D transient: total time of the time steps (s), learned+FP32 vs plain FP64",
           "value": res[main_form]["total_s"], "unit": "s", "n_gpus": 1, "steps": steps, "warmup": 0,
           "ms_per_step": res[main_form]["per_step_ms"], "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": {"learned": "f32-in-f64 + bf16 net", "fp32": "f32-in-f64", "fp64": "f64"}[main_form],
           "data": "synthetic (seeded Gaussian blob)",
           "setup_s": round(t_setup, 3),
           "config": {"workload": f"TCD3D Q1 128^3, eps 1e-4, rotating b, {'SUPG' if supg else 'Galerkin'}, "
                                  f"Crank-Nicolson tau 1e-3, {steps} steps, re-assembly every step, "
                                  f"V(2,2) SSOR omega {omega}, FGMRES(50), FP64 rel. residual 1e-10 per step",
                      "dofs": N,
                      "form": main_form, "parallelism": "1 GPU"},
           "forms": res, "speedup_vs_plain_fp64": round(t64 / res[main_form]["total_s"], 3) if t64 else None,
           "learned_training": train}
    print(json.dumps(out))
    S.close()
    ctx.close()


def train_set(args):
    """Training-set workload: learned-smoother networks trained across seeded
    coefficient fields on 64^3 Q1 (274,625 DOFs): eps = exp(G) with G a
    Gaussian random field (sigma 1, correlation length 0.1), b = (1,2,3)/sqrt(14)
    scaled to a cell Peclet number log-uniform in [10, 1000], SUPG;
    (residual, exact error) pairs from the plain cycle of the homogeneous
    problem on the three finest levels; Adam, learning rate 1e-3, batch 8.
    Reports training samples/s and convolution GFLOP/s, then the FGMRES
    solve of held-out fields with and without the networks.  Prints one JSON
    line."""
    import parmoon_amd as pa
    from parmoon_amd import prm as P
    ctx = pa.Context(0)
    S = pa.System(ctx, 2, 5, 1)
    nodes = [int(S.dims(l)[0]) for l in range(5)]
    nseeds, per_seed = args.train_set, args.train_steps
    rng = np.random.default_rng(0)
    pe = 10 ** rng.uniform(1, 3, nseeds + args.eval_seeds)
    base = dict(solver=16, restart=20, maxit=400, res_norm_min=0.0, red_factor=0.0, smoother=3, omega=1.0,
                pre=2, post=2, rel_target=1e-10)
    S.learned_init(1)
    t_train, steps, flops = 0.0, 0, 0.0
    # forward flops per sample: 3x3x3 convolutions 16 -> 16 -> 16 -> 16 -> 1 channels
    fwd = [2.0 * 27 * n * (16 * 16 + 16 * 16 + 16 * 16 + 16) for n in nodes]
    for k in range(nseeds):
        S.assemble(P.lognormal(seed=k, nref=6, peclet=pe[k]))
        info = S.learned_train(P.solver_opt(**dict(base, precision=1)), samples=8, steps=per_seed, lr=1e-3,
                               seed=k, levels=3, loss=0, data=0, batch=8)
        t_train += float(info[2])
        steps += int(info[3]) * 8  # samples: Adam steps x batch
        # forward + backward (about twice the forward) per sample
        flops += per_seed * 8 * 3.0 * (fwd[4] + fwd[3] + fwd[2])
        if (k + 1) % 16 == 0:
            print(f"[bench] train-set: {k + 1}/{nseeds} fields, {t_train:.1f}s", file=sys.stderr, flush=True)
    evals = []
    for k in range(nseeds, nseeds + args.eval_seeds):
        S.assemble(P.lognormal(seed=k, nref=6, peclet=pe[k]))
        row = {"peclet": round(float(pe[k]), 1)}
        for name, kw in (("fp64", dict(precision=0)), ("fp32", dict(precision=1)),
                         ("learned1", dict(precision=2, learned_levels=1)),
                         ("learned3", dict(precision=2, learned_levels=3)),
                         ("learned_coarse", dict(precision=2, learned_levels=3, learned_skip=1, learned_steps=2,
                                                 learned_pre=1))):
            info, rc = S.solve_resident(P.solver_opt(**dict(base, **kw)), check=False)
            row[name] = {"rc": int(rc), "iters": int(info[0]), "s": round(float(info[4]), 4)}
        evals.append(row)
        print(f"[bench] held-out field: {row}", file=sys.stderr, flush=True)
    gm = lambda key: float(np.exp(np.mean([np.log(r["fp64"]["s"] / r[key]["s"]) for r in evals])))
    out = {"metric": "learned-smoother training throughput (samples/s) on 64^3 Q1 seeded fields",
           "value": round(steps / t_train, 1), "unit": "samples/s", "n_gpus": 1, "steps": steps, "warmup": 0,
           "ms_per_step": round(t_train / steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)", "data": "synthetic (seeded Gaussian random fields)",
           "config": {"workload": f"{nseeds} lognormal fields 64^3 Q1 SUPG, Pe log-uniform 10..1000, "
                                  f"{per_seed} Adam steps of batch 8 per field and level, levels 64^3/32^3/16^3, "
                                  "(residual, exact error) pairs, lr 1e-3",
                      "parallelism": "1 GPU"},
           "conv_GFLOPs": round(flops / t_train / 1e9, 1),
           "train_seconds": round(t_train, 3),
           "note": "seconds include generating the samples with the plain cycle",
           "held_out": evals,
           "held_out_speedup_vs_fp64_geomean": {k: round(gm(k), 3)
                                                for k in ("fp32", "learned1", "learned3", "learned_coarse")}}
    print(json.dumps(out))
    S.close()
    ctx.close()


def learned_opts(args):
    """Solver options of the learned form (ignored by the other forms)."""
    return dict(learned_levels=args.learned_levels, learned_skip=args.learned_skip,
                learned_steps=args.learned_steps, learned_pre=args.learned_pre)


def learned_config(args):
    return {"levels": args.learned_levels or "all but the coarsest", "finest_levels_with_ssor": args.learned_skip,
            "steps_per_smoothing": args.learned_steps, "replaces_pre_smoothing": bool(args.learned_pre)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nref", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="learned,learned_fp16s,fp32,fp64")
    ap.add_argument("--no-offop", action="store_true", help="skip the off-operator training run")
    # learned-form defaults: the steady Q2 workload smooths the levels below
    # the finest with three network steps in place of pre- and post-smoothing;
    # the transient Q1 workload (8-colour SSOR, cheap on coarse levels) uses
    # two network steps as the finest level's post-smoother
    ap.add_argument("--learned-levels", type=int, default=None,
                    help="finest levels with a network (0: all but the coarsest)")
    ap.add_argument("--learned-skip", type=int, default=None, help="the finest of those levels keep SSOR")
    ap.add_argument("--learned-steps", type=int, default=None, help="network steps per smoothing")
    ap.add_argument("--learned-pre", type=int, default=None, help="1: the network also replaces pre-smoothing")
    ap.add_argument("--train-steps", type=int, default=None,
                    help="Adam steps per level (default: 4800 steady, 1200 transient, 16 per field for --train-set)")
    ap.add_argument("--transient", type=int, default=0, metavar="STEPS",
                    help="run the TCD3D 128^3 Q1 transient workload for STEPS time steps instead")
    ap.add_argument("--galerkin", action="store_true", help="transient: the Galerkin forms instead of SUPG")
    ap.add_argument("--omega", type=float, default=0.0, help="transient: SSOR relaxation (default 0.4 SUPG, 1 Galerkin)")
    ap.add_argument("--train-set", type=int, default=0, metavar="FIELDS",
                    help="train the learned smoother across FIELDS seeded 64^3 Q1 fields instead")
    ap.add_argument("--eval-seeds", type=int, default=4)
    args = ap.parse_args()
    if args.transient > 0:
        dflt = dict(learned_levels=1, learned_skip=0, learned_steps=2, learned_pre=0, train_steps=1200)
    elif args.train_set > 0:
        dflt = dict(learned_levels=0, learned_skip=1, learned_steps=2, learned_pre=1, train_steps=16)
    else:
        dflt = dict(learned_levels=0, learned_skip=1, learned_steps=3, learned_pre=1, train_steps=4800)
    for k, v in dflt.items():
        if getattr(args, k) is None:
            setattr(args, k, v)
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(a z-slab decomposition over several GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.transient > 0:
        transient(args)
        return
    if args.train_set > 0:
        train_set(args)
        return

    import parmoon_amd as pa
    from parmoon_amd import prm as P

    nref = args.nref
    prm = P.config2(nref=nref)
    ctx = pa.Context(0)
    # host set-up of the five levels (numbering, stencil layout, block tables)
    # and their device buffers: once per mesh, outside the steps
    t0 = time.perf_counter()
    S = pa.System(ctx, nref - 4, 5, 2)
    t_setup = time.perf_counter() - t0
    d = S.dims(-1)
    N = int(d[0])
    print(f"[bench] 2^{nref} Q2: N={N} stencil values={int(d[6])} set-up {t_setup:.3f}s", file=sys.stderr,
          flush=True)
    S.assemble(prm)

    def opt(precision):
        # res_norm_min is set per solve from the initial residual (relative 1e-10)
        return P.solver_opt(solver=16, restart=30, maxit=400, res_norm_min=0.0, red_factor=0.0,
                            smoother=3, omega=1.0, pre=2, post=2, precision=precision, rel_target=1e-10,
                            **learned_opts(args))

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    prec_of = {"fp64": 0, "fp32": 1, "learned": 2, "learned_fp16s": 4}
    train = None
    if "learned" in forms or "learned_fp16s" in forms:
        t0 = time.perf_counter()
        tinfo = S.learned_train(opt(1), samples=32, steps=args.train_steps, lr=2e-3, seed=1,
                                levels=args.learned_levels, data=1, skip=args.learned_skip)
        train = {"seconds": round(float(tinfo[2]), 3), "wall_s": round(time.perf_counter() - t0, 3),
                 "adam_steps": int(tinfo[3]), **learned_config(args),
                 "loss_first": round(float(tinfo[0]), 4), "loss_last": round(float(tinfo[1]), 4),
                 "samples_per_level": 32, "samples": "zero start, random right-hand side",
                 "loss": "scaled defect left after the correction"}
        print(f"[bench] learned smoother trained: {train}", file=sys.stderr, flush=True)

    def step(form):
        t0 = time.perf_counter()
        S.assemble(prm)
        t1 = time.perf_counter()
        info, rc = S.solve_resident(opt(prec_of[form]), check=False)
        t2 = time.perf_counter()
        if rc != 0:
            print(f"[bench] {form} solve rc={rc} info={info}", file=sys.stderr, flush=True)
        return t1 - t0, t2 - t1, info, rc

    res = {}
    for form in forms:
        for w in range(args.warmup):
            step(form)
            print(f"[bench] warmup {form} {w} done", file=sys.stderr, flush=True)
        ts_asm, ts_sol, infos = [], [], []
        for k in range(args.steps):
            ta, tsol, info, rc = step(form)
            ts_asm.append(ta)
            ts_sol.append(info[4])
            infos.append(info)
            print(f"[bench] {form} step {k}: assembly {ta:.4f}s solve {info[4]:.4f}s (wall {tsol:.4f}s) "
                  f"iters {int(info[0])} outer {int(info[3])} rel.res {info[1] / info[2]:.2e} "
                  f"fallbacks {int(info[5])}", file=sys.stderr, flush=True)
        res[form] = {"t_asm": float(np.median(ts_asm)), "t_solve": float(np.median(ts_sol)),
                     "iters": int(infos[-1][0]), "outer": int(infos[-1][3]), "fallbacks": int(infos[-1][5]),
                     "rel_res": float(infos[-1][1] / infos[-1][2]), "rc": int(rc)}

    # Rooflines on the finest level.  Algorithmic bytes are the bytes the
    # colour-sliced stencil format has to move (values are stored without
    # column indices; x counted once per launch, every launch gathers from
    # all of it):
    #   SpMV:                 stored values s_v + x (N s_v) + y of the active rows
    #   SSOR colour launch:   the colour's values + x (N s_v) + f read and x
    #                         written for the colour's rows; one launch = one
    #                         of the 54 colour applications of a sweep (the
    #                         27 colours forward and backward)
    # Durations are those of the single dispatches (events recorded by the
    # dispatch, pmx_time_kernel 10 / 11), not of back-to-back launch sequences.
    nvals = int(d[6])
    nact = int(d[1])
    ncol = int(d[7])
    kern = {}
    roof = {}
    for prec, sv in ((1, 4), (0, 8)):
        ms_spmv = S.time_kernel(11, -1, prec, reps=20)
        ms_col = S.time_kernel(10, -1, prec, reps=4)
        b_spmv = nvals * sv + N * sv + nact * sv
        b_col = nvals * sv / ncol + N * sv + 2 * nact * sv / ncol
        roof[prec] = {"spmv_ms": ms_spmv, "spmv_bytes": b_spmv, "col_ms": ms_col, "col_bytes": b_col}
    # calibration of the colour launch on the same level and with the same
    # dispatch events: a dispatch that does nothing (its workgroup counts) and
    # a plain streaming read of the colour's FP32 values alone
    cal = {"empty_dispatch_ms": S.time_kernel(12, -1, 1, reps=4),
           "stream_read_values_ms": S.time_kernel(13, -1, 1, reps=4)}
    kern["ssor_sweep_fp32_ms"] = S.time_kernel(3, -1, 1, reps=10)
    kern["ssor_sweep_fp64_ms"] = S.time_kernel(3, -1, 0, reps=10)
    for name, k in (("vcycle_fp32_ms", 7), ("restrict_fp32_ms", 6), ("prolong_fp32_ms", 5)):
        kern[name] = S.time_kernel(k, -1, 1, reps=10)
    kern["assembly_finest_ms"] = S.time_kernel(4, -1, 0, reps=3)
    # assembly of the system (every level): reset + assembly kernels + Dirichlet rhs
    t_asm_kernels = sum(S.time_kernel(4, lvl, 0, reps=3) for lvl in range(5)) * 1e-3
    conv = None
    if train is not None:
        # one network step and the forward convolutions alone on the finest
        # level that uses the network
        lv = -1 - args.learned_skip
        kern[f"learned_step_level{lv}_ms"] = S.time_kernel(8, lv, 1, reps=10)
        ms_fwd = S.time_kernel(9, lv, 1, reps=10)
        kern[f"net_forward_level{lv}_ms"] = ms_fwd
        nodes = int(S.dims(lv)[0])
        flops_fwd = 2.0 * 27 * nodes * (16 * 16 + 16 * 16 + 16 * 16 + 16)  # 16 -> 16 -> 16 -> 16 -> 1 channels
        conv = {"bound": "mfma", "kernel": "k_conv x 4 layers (network forward, BF16 in, FP32 accumulate)",
                "level_nodes": nodes, "achieved": round(flops_fwd / (ms_fwd * 1e-3) / 1e12, 2),
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(flops_fwd / (ms_fwd * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
                "flops_per_forward": flops_fwd}
    # assembly with the reference's default 64-point rule (Gauss4); the solves
    # above use the 27-point rule of INTERNAL_PROBLEM_LINEAR = 1
    prm4 = prm.copy()
    prm4[P.LINEAR] = 0.0
    S.assemble(prm4)
    asm4_finest_ms = S.time_kernel(4, -1, 0, reps=3)
    t_asm4 = sum(S.time_kernel(4, lvl, 0, reps=3) for lvl in range(5)) * 1e-3
    flops_asm4 = 428544.0 * (2 ** nref) ** 3  # Q2, 64-point rule, B^T D B model per cell

    offop = None
    if train is not None and not args.no_offop:
        # networks trained on other seeds of the Fourier potential (the same
        # steps in total, spread over three seeds), then the bench operator
        # solved with them
        seeds = (7, 11, 13)
        S.learned_init(2)
        t_off = 0.0
        for sd in seeds:
            S.assemble(P.config2(nref=nref, seed=sd))
            ti = S.learned_train(opt(1), samples=32, steps=max(1, args.train_steps // len(seeds)), lr=2e-3, seed=sd,
                                 levels=args.learned_levels, data=1, skip=args.learned_skip)
            t_off += float(ti[2])
        S.assemble(prm)
        rows = []
        for k in range(max(1, args.steps)):
            info, rc = S.solve_resident(opt(2), check=False)
            rows.append((float(info[4]), int(info[0]), int(rc), int(info[5])))
        t_o = float(np.median([r[0] for r in rows]))
        offop = {"train_seeds": list(seeds), "train_seconds": round(t_off, 3), "solve_s": round(t_o, 5),
                 "iterations": rows[-1][1], "rc": rows[-1][2], "fallbacks": rows[-1][3],
                 "speedup_vs_plain_fp64": round(res["fp64"]["t_solve"] / t_o, 3) if "fp64" in res else None}
        print(f"[bench] off-operator networks: {offop}", file=sys.stderr, flush=True)

    main_form = next(f for f in ("learned", "fp32", "fp64") if f in res)
    t_main = res[main_form]["t_solve"]
    t_asm = t_asm_kernels
    t_asm_call = res[main_form]["t_asm"]
    steps_ms = (res[main_form]["t_asm"] + res[main_form]["t_solve"]) * 1e3  # assemble call + solve
    t64 = res.get("fp64", {}).get("t_solve")
    flops_asm = 180792.0 * (2 ** nref) ** 3  # Q2, 27-point rule, B^T D B model per cell
    r32 = roof[1]
    col_gbs = r32["col_bytes"] / (r32["col_ms"] * 1e-3) / 1e9
    spmv_gbs = r32["spmv_bytes"] / (r32["spmv_ms"] * 1e-3) / 1e9
    asm_tf = flops_asm / (kern["assembly_finest_ms"] * 1e-3) / 1e12
    asm4_tf = flops_asm4 / (asm4_finest_ms * 1e-3) / 1e12
    out = {
        "metric": "time-to-1e-10 FP64 residual (s), learned+FP32 vs plain FP64 GMG; assembly DOF/s",
        "value": round(t_main, 5),
        "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(steps_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"learned": "f32-in-f64 + bf16 net", "fp32": "f32-in-f64", "fp64": "f64"}[main_form],
        "data": "synthetic (seeded Fourier convection field)",
        "config": {"workload": f"CD3D SUPG Q2 {2 ** nref}^3, eps 1e-4, curl-Fourier b, 5-level GMG V(2,2) "
                               f"multicolour SSOR, FGMRES(30), FP64 rel. residual 1e-10", "dofs": N,
                   "form": main_form, "parallelism": "1 GPU"},
        "plain_fp64_time_s": t64,
        "fp32_time_s": res.get("fp32", {}).get("t_solve"),
        "learned_time_s": res.get("learned", {}).get("t_solve"),
        "learned_fp16_smoother_time_s": res.get("learned_fp16s", {}).get("t_solve"),
        "speedup_vs_plain_fp64": round(t64 / t_main, 3) if t64 else None,
        "speedup_learned_fp16_smoother_vs_plain_fp64": (round(t64 / res["learned_fp16s"]["t_solve"], 3)
                                                        if t64 and "learned_fp16s" in res else None),
        # training counted once, with one solve: the whole cost of the learned
        # form for a single right-hand side
        "learned_with_training_s": (round(train["seconds"] + t_main, 4) if train and main_form == "learned"
                                    else None),
        "learned_off_operator": offop,
        "learned_training": train,
        "setup_s": round(t_setup, 3),
        "assembly_dofs_per_s": round(N / t_asm, 1),
        "assembly_s": round(t_asm, 5),
        "assemble_call_s": round(t_asm_call, 5),
        "assembly_gauss4": {"dofs_per_s": round(N / t_asm4, 1), "s": round(t_asm4, 5),
                            "finest_ms": round(asm4_finest_ms, 4)},
        "iterations": {f: res[f]["iters"] for f in res},
        "fallbacks": {f: res[f]["fallbacks"] for f in res},
        "final_rel_residual": {f: res[f]["rel_res"] for f in res},
        "kernels_ms": {"spmv_fp32_dispatch": round(r32["spmv_ms"], 4), "spmv_fp64_dispatch": round(roof[0]["spmv_ms"], 4),
                       "ssor_colour_fp32_dispatch": round(r32["col_ms"], 4),
                       "ssor_colour_fp64_dispatch": round(roof[0]["col_ms"], 4),
                       **{k: round(v, 4) for k, v in kern.items()}},
        # the dominant kernel of the timed solve: the colour launches of the
        # finest level's FP32 SSOR sweeps
        "roofline": {"bound": "hbm", "kernel": "k_rows<float,3,1> (one SSOR colour launch, finest level)",
                     "achieved": round(col_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(col_gbs / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_SSOR_LAUNCH,
                     "algorithmic_bytes": int(r32["col_bytes"]), "dispatch_ms": round(r32["col_ms"], 5),
                     "calibration": {"empty_dispatch_ms": round(cal["empty_dispatch_ms"], 5),
                                     "stream_read_of_the_values_ms": round(cal["stream_read_values_ms"], 5),
                                     "launch_over_stream_read": round(r32["col_ms"] / cal["stream_read_values_ms"],
                                                                      3)}},
        "roofline_spmv": {"bound": "hbm", "kernel": "k_rows<float,0,1> (stencil SpMV, finest level)",
                          "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_SPMV,
                          "algorithmic_bytes": int(r32["spmv_bytes"]), "dispatch_ms": round(r32["spmv_ms"], 5),
                          "fp64_GBs": round(roof[0]["spmv_bytes"] / (roof[0]["spmv_ms"] * 1e-3) / 1e9, 1)},
        # Q2 assembly of the finest level against the FP64 peak (B^T D B model:
        # 47.4 GFLOP with the 27-point rule, 112.3 GFLOP with the 64-point rule)
        "roofline_assembly": {"bound": "mfma", "kernel": "k_qp_data + k_cell_mfma + k_assemble_kc (finest level)",
                              "achieved": round(asm_tf, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(asm_tf / FP64_PEAK_TFLOPS, 4),
                              "gauss4_achieved": round(asm4_tf, 2),
                              "gauss4_frac": round(asm4_tf / FP64_PEAK_TFLOPS, 4)},
        "roofline_conv": conv,
    }
    if not args.no_cpu:
        try:
            kind = "reference" if os.path.exists(REF_SOLVE) else "port"
            c = cpu_baseline_reference() if kind == "reference" else cpu_baseline()
            who = ("ParMooN's own code (oracle/_ref/ref_solve, compiled from the reference sources)"
                   if kind == "reference" else "oracle (C++ restatement of the reference algorithms)")
            model, ncpu = cpu_info()
            its64 = res.get("fp64", {}).get("iters")
            # time per iteration and DOF of the sample, times the DOFs and the
            # plain FP64 iteration count of the full workload (the same method:
            # FGMRES(30), V(2,2) SSOR; the oracle sweeps lexicographically)
            scale_n = N / c["N"]
            scale_it = (its64 / c["iters"]) if its64 else 1.0
            out["cpu_baseline"] = {
                "value": round(c["t_solve"] * scale_n * scale_it, 2), "unit": "s", "cores": 1, "kind": kind,
                "sample": f"{who}, one core, on the same workload at 32^3 Q2 ({c['N']} DOFs): solve to 1e-10 in "
                          f"{c['iters']} FGMRES iterations, {c['t_solve']:.2f} s; value = that time x DOF ratio "
                          f"{scale_n:.2f} x iteration ratio {scale_it:.3f} (GPU FP64 iterations at {2 ** nref}^3)",
                "sample_solve_s": round(c["t_solve"], 3), "sample_iterations": c["iters"],
                "sample_setup_s": round(c["t_setup"], 3), "sample_assembly_s": round(c["t_assembly_all"], 3),
                "assembly_dofs_per_s": round(c["N"] / c["t_assembly_all"], 1),
                "cpu_model": model, "nproc": ncpu}
        except Exception as e:  # the CPU leg never hides the GPU line
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
