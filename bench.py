#!/usr/bin/env python3
"""LBM-PSM throughput on one MI355X.

Headline workload: the D3Q19 BGK-PSM fluidised-bed column, 1024 x 256 x 256
nodes (x vertical, `boundary f p p`), `lbm-psm-bc xFlow` (Zou-He velocity
inlet U = 0.005 at x = 0, density outlet rho = 1, and velocity faces on the
y/z planes as the reference's xFlow applies them), gravity Fx = -2e-6,
tau = 0.7, 20000 spheres d = 12 seeded by RSA over the column (phi = 0.27),
moving linearly (v ~ U(+-1e-3), omega ~ U(+-1e-4)) so that every step maps
new positions.  Secondary workload: D3Q19 192^3 shear suspension (y walls at
+-0.01, 400 moving spheres d = 16).

One step = one LBM step of `fix lbm-psm`: map the atom list (locals + ghost
images) onto the lattice, fused pull-stream + Zou-He + BGK-PSM collide with
Guo forcing, and the per-particle force/torque/stresslet reduction.  Every
step's list is staged in device memory before the timed region (`value`,
FP64 storage); the same steps with a host upload, a result download and a
synchronisation per step, as the fix runs them, are `host_inclusive`.
FP32 and FP16 storage are reported under `variants` with their drag and
torque error against FP64 after the same steps.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--storage f64|f32|f16]
                  [--workload bed|shear]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lammps-lbm-psm_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
Q = 19
METRIC = "MLUPS + % HBM roofline, D3Q19 PSM: FP64 vs FP32-storage at stated drag error"
BYTES_PER_ELEM = {"f64": 8, "f32": 4, "f16": 2}
PROF_KERNELS = ("collide_stream", "map_bin", "map_fraction", "map_particles", "hydro_reduce")
KERNEL_NAME = {"f64": "k_collide_f64<3,19,false", "f32": "k_collide_lowp<3,19,1,false",
               "f16": "k_collide_lowp<3,19,2,false"}


def kernel_name(storage, F, sz):
    """The collide kernel the library launches: formed for a body force along
    one lattice axis (that axis), else for any force (-1); then whether the
    stored z rows are a whole number of 64-node blocks (split addressing)."""
    nz = [a for a in range(3) if F[a] != 0.0]
    aln = "true" if sz % 64 == 0 else "false"
    return f"{KERNEL_NAME[storage]},{nz[0] if len(nz) == 1 else -1},{aln}>"


class Workload:
    def __init__(self, name):
        from lbmpsm import workloads as W
        self.name = name
        if name == "bed":
            c = W.BED
            self.box, self.atoms = W.fluidised_bed()
            self.n, self.per, self.tau, self.F = c["n"], c["periodic"], c["tau"], c["F"]
            self.u_lb, self.bc, self.radius = c["u_inlet"], c["bc"], c["radius"]
            self.describe = ("D3Q19 BGK-PSM fluidised-bed column 1024x256x256 (x vertical), "
                             "xFlow (Zou-He inlet U=0.005, outlet rho=1), Fx=-2e-6, tau=0.7, "
                             "20000 moving spheres d=12 (seeded RSA, phi=0.27)")
        else:
            c = W.SHEAR
            self.box, self.atoms = W.shear_suspension()
            self.n, self.per, self.tau, self.F = c["n"], c["periodic"], c["tau"], (0.0, 0.0, 0.0)
            self.u_lb, self.bc, self.radius = c["u_wall"], c["bc"], c["radius"]
            self.describe = ("D3Q19 BGK-PSM 192^3, shear (Zou-He walls y, U=0.01), "
                             "400 moving spheres d=16 (seeded RSA), tau=0.7")
        self.nodes = self.n[0] * self.n[1] * self.n[2]
        self.cutoff = self.radius + 4.0
        self.bc_id = {"shear": 1, "xFlow": 2}[self.bc]

    def lists(self, nsteps):
        from lbmpsm import workloads as W
        return W.atom_lists(self.box, self.atoms, nsteps, self.cutoff)

    def lattice(self, storage):
        import lbmpsm
        cfg = lbmpsm.lattice_units_config(3, self.n, self.per, self.tau, F_lbm=self.F,
                                          storage=storage, u_lb=self.u_lb)
        lat = lbmpsm.Lattice(cfg)
        lat.set_bc(self.bc)
        return lat


def stage_lists(lists):
    """All per-step atom lists in device memory (one concatenated block per
    array), so the timed loop starts with its inputs resident in HBM."""
    from lbmpsm.devmem import DeviceArray
    off = np.cumsum([0] + [L.n for L in lists])
    cat = lambda f, dt: DeviceArray(np.ascontiguousarray(np.concatenate([f(L) for L in lists]),
                                                         dtype=dt))
    block = dict(tag=cat(lambda L: L.tag, np.int32), x=cat(lambda L: L.x, np.float64),
                 v=cat(lambda L: L.v, np.float64), w=cat(lambda L: L.omega, np.float64),
                 r=cat(lambda L: L.radius, np.float64))
    rmax = max(float(np.max(L.radius)) for L in lists)
    return block, off, rmax


def run_gpu(wl, storage, staged, warmup, steps):
    """Device-resident run: lists already in HBM, map + collide + faces +
    hydro per step with no host transfer inside the timed region."""
    block, off, rmax = staged
    lat = wl.lattice(storage)
    p = {k: t.address for k, t in block.items()}

    def one(s):
        o, n = int(off[s]), int(off[s + 1] - off[s])
        lat.set_particles_device(n, p["tag"] + 4 * o, p["x"] + 24 * o, p["v"] + 24 * o,
                                 p["w"] + 24 * o, p["r"] + 8 * o, rmax)
        lat.step()

    for s in range(warmup):
        one(s)
    lat.sync()
    # timed region: HIP events on the library's stream around the bulk
    # collide only (the kernel the roofline figure is for)
    lat.profile("collide")
    lat.profile_reset()
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        one(s)
    lat.sync()
    t1 = time.perf_counter()
    collide = lat.profile_read("collide_stream")
    F, T, S = lat.hydro()
    # per-kernel breakdown from a separate, untimed pass over the warm-up lists
    lat.profile(True)
    lat.profile_reset()
    for s in range(min(warmup, 10)):
        one(s)
    lat.sync()
    prof = {k: lat.profile_read(k) for k in PROF_KERNELS}
    prof["collide_stream"] = collide
    lat.profile(False)
    info = lat.info
    lat.close()
    return dict(el=t1 - t0, F=F, T=T, S=S, prof=prof, info=info)


def run_gpu_host(wl, storage, lists, warmup, steps):
    """As `fix lbm-psm` drives it: host lists uploaded, results downloaded
    and the stream synchronised every step."""
    lat = wl.lattice(storage)
    for s in range(warmup):
        L = lists[s]
        lat.set_particles(L.tag, L.x, L.v, L.omega, L.radius)
        lat.step()
        lat.hydro()
    lat.sync()
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        L = lists[s]
        lat.set_particles(L.tag, L.x, L.v, L.omega, L.radius)
        lat.step()
        lat.hydro()
    t1 = time.perf_counter()
    lat.close()
    return t1 - t0


def roofline(wl, storage, res):
    prof, info = res["prof"], res["info"]
    ms_c, n_c = prof["collide_stream"]
    bytes_per_update = 2 * Q * BYTES_PER_ELEM[storage]
    avg_ms = ms_c / max(n_c, 1)
    # the collide updates every stored node (a wall axis keeps its two halo
    # planes, whose populations the pull reads and writes too)
    nodes = int(info.stored_nodes)
    achieved = bytes_per_update * nodes / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kernel_name(storage, wl.F, info.sz), "bytes_per_update": bytes_per_update,
            "nodes_per_launch": nodes, "avg_launch_ms": round(avg_ms, 5)}
    # HBM bytes per launch of the dominant kernel from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy kernel)
    tpath = os.path.join(ROOT, "profiles", f"collide_{storage}_{wl.name}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        roof["traffic"] = tr.get("hbm_bytes_per_launch")
        roof["traffic_source"] = os.path.relpath(tpath, ROOT)
    return roof


def errors_vs_f64(wl, lst, res, ref):
    """Max over spheres of |X - X_f64| / max |X_f64| for the owner-summed
    drag and torque (what the fix hands to LAMMPS)."""
    from lbmpsm import harness as H
    F64, T64 = H.reduce_to_owners(lst, ref["F"], ref["T"])
    F, T = H.reduce_to_owners(lst, res["F"], res["T"])
    eF = float(np.max(np.linalg.norm(F - F64, axis=1)) / np.max(np.linalg.norm(F64, axis=1)))
    eT = float(np.max(np.linalg.norm(T - T64, axis=1)) / np.max(np.linalg.norm(T64, axis=1)))
    return eF, eT


def cpu_baseline(wl, lists, steps_all=2):
    """The CPU oracle (reference algorithm and layout, OpenMP over lattice
    planes and over the nodes of a particle's zone): `steps_all` full steps
    of the workload on all host threads after one untimed step, and one step
    on one thread of a 128-plane section of the same column (its spheres)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lbmpsm_oracle as O
    from lbmpsm import harness as H
    nthreads = O.threads()
    ora = O.Oracle(3, wl.n, wl.per, wl.tau, F_lbm=wl.F, unit=O.lattice_units(wl.u_lb))
    ora.fix_step(lists[0], bc_keyword=wl.bc_id)
    t0 = time.perf_counter()
    for s in range(1, steps_all + 1):
        ora.fix_step(lists[s], bc_keyword=wl.bc_id)
    t1 = time.perf_counter()
    ora.close()
    all_mlups = wl.nodes * steps_all / (t1 - t0) / 1e6
    # one thread: a section of x-planes [0, 128) of the column with the
    # spheres whose zones reach into it (wall axis x keeps its faces)
    sec = 128 if wl.n[0] > 128 else wl.n[0]
    n1 = (sec, wl.n[1], wl.n[2])
    box1 = H.Box(lo=np.zeros(3), prd=np.array([sec - 1.0 if not wl.per[0] else float(sec),
                                              wl.box.prd[1], wl.box.prd[2]]),
                 periodic=wl.per, dim=3)
    keep = wl.atoms.x[:, 0] < sec + wl.cutoff
    a1 = H.Atoms(tag=wl.atoms.tag[keep], x=wl.atoms.x[keep], v=wl.atoms.v[keep],
                 omega=wl.atoms.omega[keep], radius=wl.atoms.radius[keep])
    l1 = [H.build_list(a1, box1, wl.cutoff) for _ in range(2)]
    O.set_threads(1)
    ora = O.Oracle(3, n1, wl.per, wl.tau, F_lbm=wl.F, unit=O.lattice_units(wl.u_lb))
    ora.fix_step(l1[0], bc_keyword=wl.bc_id)
    t0 = time.perf_counter()
    ora.fix_step(l1[1], bc_keyword=wl.bc_id)
    t1 = time.perf_counter()
    ora.close()
    O.set_threads(nthreads)
    one_mlups = n1[0] * n1[1] * n1[2] / (t1 - t0) / 1e6
    return {"value": round(all_mlups, 3), "unit": "MLUPS", "cores": nthreads, "kind": "port",
            "sample": f"{steps_all} full FP64 steps of the same workload ({wl.nodes} nodes, "
                      f"{lists[0].n} list entries) after 1 untimed step, {nthreads} OpenMP "
                      f"threads",
            "one_core": {"value": round(one_mlups, 3), "unit": "MLUPS", "cores": 1,
                         "sample": f"1 FP64 step (after 1 untimed) of x-planes [0, {sec}) of "
                                   f"the same column with its {int(keep.sum())} spheres, "
                                   f"1 thread"}}


def reference_benchmark_deck():
    """The reference's own published benchmark (benchmark/in.lbmpsm_1:8-40):
    a fixed sphere of d = 0.0005 m at the centre of a 0.005 m cube,
    `boundary p f p`, `fix lbm-psm ... Nlc 31 lc 0.0005 rho 1000 nu 1e-4 Re 0.1
    tau 0.7` (300 x 301 x 300 = 27,090,000 nodes), `lbm-psm-bc shear`, 100
    steps, driven as the fix drives it (host list, step, F/T/S download every
    step).  The published quantity is LAMMPS's total wall time on a 2 x Xeon
    Silver 4114 workstation (benchmark/benchmark-results.dat:2-7, copied into
    tests/golden/reference_verification.json); here the lattice set-up and
    the 100 steps are timed on one GPU, FP64, without the deck's VTI output."""
    import lbmpsm
    from lbmpsm import harness as H
    box_l, r = 0.005, 0.00025
    cfg = lbmpsm.config_from_fix(3, 31, 0.0005, 1000.0, 1e-4, 0.1, 0.7, (0.0, 0.0, 0.0),
                                 (0.0, 0.0, 0.0), (box_l, box_l, box_l), (1, 0, 1))
    box = H.Box(lo=np.zeros(3), prd=np.array([box_l] * 3), periodic=(1, 0, 1), dim=3)
    atoms = H.Atoms(tag=np.array([1], np.int32), x=np.array([[0.0025] * 3]), v=np.zeros((1, 3)),
                    omega=np.zeros((1, 3)), radius=np.array([r]))
    lst = H.build_list(atoms, box, cutoff=0.001)
    t0 = time.perf_counter()
    lat = lbmpsm.Lattice(cfg)
    lat.set_bc("shear")
    lat.sync()
    t1 = time.perf_counter()
    for _ in range(100):
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
        lat.hydro()
    t2 = time.perf_counter()
    nodes = int(np.prod(lat.shape))
    lat.close()
    with open(os.path.join(ROOT, "tests", "golden", "reference_verification.json")) as fh:
        pub = json.load(fh)["benchmark_wall_time_s"]
    return {"deck": "benchmark/in.lbmpsm_1 (one fixed sphere in shear, Nlc 31, 100 steps)",
            "nodes": nodes, "setup_s": round(t1 - t0, 3), "steps_s": round(t2 - t1, 3),
            "MLUPS_steps": round(nodes * 100 / (t2 - t1) / 1e6, 1),
            "published_wall_s": {f"{int(c)} cores": w for c, w in pub["rows"]},
            "published_source": pub["source"],
            "note": "published: LAMMPS total wall time incl. set-up and VTI output, CPU"}


def measure(wl, storages, warmup, steps, host=True):
    lists = wl.lists(warmup + steps)
    staged = stage_lists(lists)
    res = {}
    for st in storages:
        print(f"[bench] {wl.name} storage {st}: {warmup} warm-up + {steps} timed steps",
              file=sys.stderr, flush=True)
        res[st] = run_gpu(wl, st, staged, warmup, steps)
    host_el = None
    if host:
        print(f"[bench] {wl.name} storage {storages[0]}, host lists and per-step download",
              file=sys.stderr, flush=True)
        host_el = run_gpu_host(wl, storages[0], lists, warmup, steps)
    return lists, res, host_el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--storage", default="f64", choices=["f64", "f32", "f16"])
    ap.add_argument("--workload", default="bed", choices=["bed", "shear"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the lattice would shard by slabs along one axis, exchanging the 5 "
              "populations that cross each face every step).", file=sys.stderr)
        sys.exit(2)

    wl = Workload(args.workload)
    storages = [args.storage] + ([] if args.no_variants else
                                 [s for s in ("f64", "f32", "f16") if s != args.storage])
    lists, res, host_el = measure(wl, storages, args.warmup, args.steps,
                                  host=not args.no_variants)
    main_res = res[args.storage]
    el = main_res["el"]
    ms_step = el / args.steps * 1e3
    mlups = wl.nodes * args.steps / el / 1e6
    roof = roofline(wl, args.storage, main_res)
    from lbmpsm import probe
    print("[bench] copy ceiling", file=sys.stderr, flush=True)
    copy_gbs, copy_form, copy_all = probe.copy_ceiling()
    roof["copy_ceiling"] = round(copy_gbs, 1)
    roof["copy_form"] = probe.FORMS[copy_form]
    roof["copy_forms"] = {probe.FORMS[f]: round(v, 1) for f, v in copy_all.items()}
    roof["frac_of_copy"] = round(roof["achieved"] / copy_gbs, 4)
    # the collide's own access pattern with no arithmetic, at the stored box
    info = main_res["info"]
    if info.sz % 64 == 0:
        print("[bench] pull pattern", file=sys.stderr, flush=True)
        pat_ms = probe.pull_pattern(info.sx, info.sy, info.sz, BYTES_PER_ELEM[args.storage])
        roof["pattern_ms"] = round(pat_ms, 5)
        roof["pattern_gbs"] = round(roof["bytes_per_update"] * roof["nodes_per_launch"] /
                                    (pat_ms * 1e-3) / 1e9, 1)
        roof["frac_of_pattern"] = round(pat_ms / roof["avg_launch_ms"], 4)
    kernels = {k: {"avg_ms": round(v[0] / max(v[1], 1), 5), "launches": v[1]}
               for k, v in main_res["prof"].items()}
    variants = {}
    for st, r in res.items():
        d = {"MLUPS": round(wl.nodes * args.steps / r["el"] / 1e6, 1),
             "ms_per_step": round(r["el"] / args.steps * 1e3, 4),
             "collide_ms": round(r["prof"]["collide_stream"][0] /
                                 max(r["prof"]["collide_stream"][1], 1), 5),
             "collide_frac_of_peak": roofline(wl, st, r)["frac"]}
        if "f64" in res and st != "f64":
            eF, eT = errors_vs_f64(wl, lists[-1], r, res["f64"])
            d["drag_rel_err_vs_f64"] = eF
            d["torque_rel_err_vs_f64"] = eT
        variants[st] = d
    if "f64" in res and "f32" in res:
        variants["f32_over_f64"] = round(res["f64"]["el"] / res["f32"]["el"], 3)

    secondary = None
    if not args.no_secondary:
        other = Workload("shear" if wl.name == "bed" else "bed")
        _, res2, _ = measure(other, ["f64", "f32"], args.warmup, args.steps, host=False)
        secondary = {"workload": other.describe}
        for st, r in res2.items():
            secondary[st] = {"MLUPS": round(other.nodes * args.steps / r["el"] / 1e6, 1),
                             "ms_per_step": round(r["el"] / args.steps * 1e3, 4),
                             "roofline": roofline(other, st, r)}
        print("[bench] the reference's benchmark deck", file=sys.stderr, flush=True)
        secondary["reference_benchmark"] = reference_benchmark_deck()

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl, lists)

    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic",
        "config": {"workload": wl.describe, "nodes": wl.nodes, "spheres": int(wl.atoms.n),
                   "list_entries": int(lists[-1].n), "storage": args.storage},
        "roofline": roof,
        "kernels": kernels,
        "variants": variants,
        "host_inclusive": None if host_el is None else {
            "MLUPS": round(wl.nodes * args.steps / host_el / 1e6, 1),
            "ms_per_step": round(host_el / args.steps * 1e3, 4),
            "note": "per step: host list upload, map, step, F/T/S download, sync (as the fix)"},
        "secondary": secondary,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
