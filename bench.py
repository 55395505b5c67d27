#!/usr/bin/env python3
"""bench.py -- voxel-to-mesh throughput on one MI355X.

Workload (one "step"): the 512x512x512 float32 volume of the large-pipeline
workload (the 128^3 head generator scaled 4x, noise sigma 8), already in HBM:

  1. dark / mid / bright multi-Otsu isolevels from GPU histograms
     (setThreshold modes 1, 2, 3, src/isolevel.c:245-277);
  2. for each isolevel: meshify() with pre-smooth, largest cluster and bubble
     fill, Lewiner MC33 marching cubes, vertex welding and degenerate removal
     (src/meshify.c:286-389) -- the pre-smooth is computed once and shared;
  3. 100 HC smoothing iterations per mesh (alpha 0.1, beta 0.5, locked edges,
     src/quadric.c:343-394).

Metric: GVoxel/s = 512^3 voxels per step / step time.  The quadric
reduction of the workload stays the reference's CPU code and is not part of
the step.

The second half of the metric, batched atlas vs per-label loop, is measured
on the 256^3 / 370-label Voronoi atlas (`atlas` object of the JSON line):
every label binarised at 0.5, pre-smoothed and meshed (classic marching
cubes, the CLI default), once by the batched path and once by the
reference's per-label loop run on the same GPU; both give the same meshes.

Also printed in the same JSON line:
  roofline      the dominant kernel's algorithmic bytes per launch / its mean
                launch time (HIP events on the library's stream, during the
                timed steps), against the 8 TB/s HBM peak.  The dominant
                kernel is k_hc_ring on the largest mesh (stage "hc_ring"),
                whose 100 iterations run alone on the GPU; the smaller
                meshes' iterations overlap its preparation and are timed as
                "hc_ring_side";
  cpu_baseline  the reference compiled from its own sources, timed on this
                host (1 core) on a bounded sample of the same workload, with
                its stage timers (-v 1) and its quadric reduction at 0.25
                beside the library's quadric_simplify_mesh on the same meshes;
  pcie_inclusive  the same step fed from, and returning to, host memory;
  host_cpu      the host CPU model and the CPUs this process may use;
  atlas.cpu_loop  the reference CLI's -a loop (OpenMP build) at 1 thread and
                at every CPU, on a slab of the atlas.
The roofline's `traffic` is read from profiles/traffic_<kernel>.txt
(rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE of the same kernel, calibrated
against a copy kernel): a committed measurement, not one of this run.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 512
HC_ITERS = 100
MODES = [1, 2, 3]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


REF_DIR = os.path.join(ROOT, 'oracle', '_ref')
STAGES = ('pre-smooth', 'voxel clustering', 'marching cubes', 'vertex welding', 'remove degenerate triangles',
          'post-smooth')


def host_cpu():
    """CPU model and the logical CPUs this process may use."""
    model = 'unknown'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return {'model': model, 'logical_cpus': len(os.sched_getaffinity(0))}


def ref_exe(name):
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        raise SystemExit('bench.py: the reference build %s is missing (make -C oracle ref); '
                         'run with --no-cpu to skip the CPU baseline' % path)
    return path


def run_verbose(args, env=None):
    """Run a reference CLI with -v 1; returns {stage: total ms} from its
    stage lines ('<stage>...: N ms', src/meshify.c:303-380, src/nii2mesh.c:330-355)."""
    import re
    import subprocess
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        raise RuntimeError('%s failed: %s' % (' '.join(args), r.stdout[-500:]))
    ms = {}
    for line in r.stdout.splitlines():
        m = re.match(r'^([a-z][a-z -]*?)(?: \(.*\)| [0-9].*)?: (\d+) ms$', line.strip())
        if m:
            ms[m.group(1)] = ms.get(m.group(1), 0) + int(m.group(2))
    return ms, r.stdout


def cpu_baseline(vol, isos, z0=224, nz=64):
    """The reference CLI built from its own sources (oracle/_ref, Lewiner
    build) on a z-slab of the same volume, on this host, one thread (the
    path is single-threaded): setThreshold for the three modes (timed
    through the reference's own setThreshold), then per isolevel
    `nii2mesh slab.nii -i <iso> -l 1 -b 1 -s 100 -r 0.25 -v 1`; the hot-path
    time is the sum of its pre-smooth, clustering, marching-cubes, weld,
    degenerate and post-smooth stage lines (load and save excluded), the
    quadric reduction at 0.25 is reported beside it."""
    import tempfile
    from tests import support  # checker bindings; the checker is only timed here
    exe = ref_exe('nii2mesh_lewiner')
    lib = support.Ref(C.CDLL(os.path.join(REF_DIR, 'libn2mref_lewiner.so')))
    slab = np.ascontiguousarray(vol[z0:z0 + nz])
    t0 = time.perf_counter()
    for m in MODES:
        lib.setThreshold(slab, m)
    thr_ms = (time.perf_counter() - t0) * 1e3
    tmp = tempfile.mkdtemp(dir='/dev/shm' if os.path.isdir('/dev/shm') else None)
    stages = {k: 0 for k in STAGES}
    quadric_ms = 0
    try:
        nii = os.path.join(tmp, 'slab.nii')
        support.write_nifti(nii, slab, 16)
        for iso in isos:
            ms, _ = run_verbose([exe, nii, '-i', '%.9g' % iso, '-l', '1', '-b', '1', '-s', str(HC_ITERS),
                                 '-r', '0.25', '-v', '1', os.path.join(tmp, 'out.mz3')])
            for k in STAGES:
                stages[k] += ms.get(k, 0)
            quadric_ms += ms.get('simplify vertices', 0)
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    hot_ms = thr_ms + sum(stages.values())
    # the library's quadric_simplify_mesh (first pass on the GPU) on the same
    # slab meshes, meshed and smoothed by the library (bit-identical to the
    # reference's)
    import nii2mesh_amd as n2m
    from nii2mesh_amd import abi
    q = abi.bind_quadric(n2m.load())
    glib = n2m.lib(n2m.CUBES_LEWINER)
    lib_ms = 0.0
    for iso in isos:
        rc, t, p = glib.meshify(slab.copy(), iso, True, True, True)
        if rc != 0:
            continue
        glib.laplacian_smoothHC(p, t, 0.1, 0.5, HC_ITERS, True)
        target = int(np.floor(np.float32(len(t)) * np.float32(0.25) + 0.5))  # round((float)ntri * 0.25)
        t0 = time.perf_counter()
        q(p, t, target, 7.0)
        lib_ms += (time.perf_counter() - t0) * 1e3
    return {'value': slab.size / (hot_ms * 1e-3) / 1e9, 'unit': 'GVoxel/s', 'cores': 1, 'kind': 'reference',
            'sample': '512x512x%d z-slab [%d,%d) of the same volume: setThreshold x3 + 3 isolevels x nii2mesh '
                      '(Lewiner build) -l 1 -b 1 -s %d, stage lines of -v 1, load/save excluded: %.1f s'
                      % (nz, z0, z0 + nz, HC_ITERS, hot_ms / 1e3),
            'stages_ms': dict(stages, setThreshold=round(thr_ms, 1)),
            'quadric_r0.25_ms': quadric_ms,
            'library_quadric_r0.25_ms': round(lib_ms, 1)}


def cpu_atlas(vol, z0=112, nz=32):
    """The reference CLI's atlas loop (-a 1, src/nii2mesh.c:492-573; classic
    build with OpenMP) on a z-slab of the 256^3 atlas, with one thread and
    with every CPU this process may use; -r 1 -q 0, load and save excluded."""
    import tempfile
    import shutil
    from tests import support
    exe = ref_exe('nii2mesh_classic_omp')
    slab = np.ascontiguousarray(vol[z0:z0 + nz])
    tmp = tempfile.mkdtemp(dir='/dev/shm' if os.path.isdir('/dev/shm') else None)
    out = {'sample': '256x256x%d z-slab [%d,%d) of the 256^3 atlas, labels 1..%d, nii2mesh -a 1 -r 1 -q 0'
                     % (nz, z0, z0 + nz, int(slab.max()))}
    try:
        nii = os.path.join(tmp, 'atlas.nii')
        support.write_nifti(nii, slab, 16)
        for n in (1, len(os.sched_getaffinity(0))):
            env = dict(os.environ, OMP_NUM_THREADS=str(n))
            t0 = time.perf_counter()
            ms, text = run_verbose([exe, nii, '-a', '1', '-r', '1', '-q', '0', '-v', '1',
                                    os.path.join(tmp, 'lab.mz3')], env)
            wall = (time.perf_counter() - t0) * 1e3
            io = ms.get('load from disk', 0) + ms.get('save to disk', 0)
            key = 'threads_%d_ms' % n
            # with several threads the per-label save lines overlap: subtract
            # them only in the one-thread run, where they are sequential
            out[key] = round(wall - io if n == 1 else wall - ms.get('load from disk', 0), 1)
            out['labels_meshed'] = text.count('save to disk')
        out['threads'] = len(os.sched_getaffinity(0))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def pcie_inclusive(torch, vol, step, ctx):
    """The same step fed from and returning to host memory: the volume
    copied in from pageable memory, the three meshes copied out."""
    host = torch.from_numpy(vol)
    d = torch.empty(vol.shape, dtype=torch.float32, device='cuda')
    best = None
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.copy_(host)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step(d)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for k in range(3):
            ctx.mesh(k)
        t3 = time.perf_counter()
        r = ((t1 - t0) * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3)
        if best is None or sum(r) < sum(best):
            best = r
    del d
    total = sum(best)
    return {'value': round(vol.size / (total * 1e-3) / 1e9, 4), 'unit': 'GVoxel/s', 'ms_per_step': round(total, 3),
            'h2d_ms': round(best[0], 3), 'step_ms': round(best[1], 3), 'd2h_ms': round(best[2], 3),
            'note': 'volume from pageable host memory, meshes to host arrays'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-atlas', action='store_true', help='skip the batched-atlas leg')
    ap.add_argument('--calib-copy', action='store_true',
                    help='also copy the 512 MiB volume once with a blit and once with an elementwise kernel '
                         '(known-size kernels that calibrate the FETCH_SIZE / WRITE_SIZE counters)')
    ap.add_argument('--atlas-only', action='store_true', help='run only the batched-atlas leg')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU; --gpus %d is not supported (the path would '
              'shard as z-slabs with halo exchange, not implemented here)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import nii2mesh_amd as n2m
    from nii2mesh_amd import synth

    if args.atlas_only:
        print(json.dumps({'atlas': atlas_leg(torch, n2m, synth)[0]}), flush=True)
        return
    t = time.perf_counter()
    vol = synth.large_f32()
    log('generated 512^3 volume in %.1f s' % (time.perf_counter() - t))
    d = torch.from_numpy(vol).to('cuda')
    torch.cuda.synchronize()
    if args.calib_copy:
        e = torch.empty_like(d)
        e.copy_(d)
        torch.mul(d, 1.0, out=e)
        torch.cuda.synchronize()
        del e
    nvox = vol.size
    dims = (N, N, N)
    ctx = n2m.DeviceContext()
    prm = ctx.params(0.0, cubes=n2m.CUBES_LEWINER, preSmooth=True, onlyLargest=True, fillBubbles=True)
    L = ctx.L
    L.n2m_ctx_set_thresholds.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_float)]
    modes = (C.c_int * 3)(*MODES)
    isos_c = (C.c_float * 3)()

    def step(dv=None):
        dv = d if dv is None else dv
        if L.n2m_ctx_set_thresholds(ctx.h, C.c_void_p(dv.data_ptr()), nvox, 3, modes, isos_c) != 0:
            raise RuntimeError(n2m.last_error())
        isos = [float(v) for v in isos_c]
        if ctx.pipeline(dv.data_ptr(), dims, prm, isos, 0.1, 0.5, HC_ITERS) != 0:
            raise RuntimeError('pipeline failed: ' + n2m.last_error())
        return isos

    # warmup; the last warmup step times every stage (the per-kernel table)
    stats = {}
    for i in range(args.warmup):
        last = i == args.warmup - 1
        if last:
            torch.cuda.synchronize()
            ctx.timing(True)
            ctx.reset_timing()
        isos = step()
        torch.cuda.synchronize()
        if last:
            stats = ctx.kernel_stats()
        log('warmup %d done, isolevels %s' % (i, isos))
    tot_ms = sum(v[0] for v in stats.values())
    tot_bytes = sum(v[2] for v in stats.values())
    for k, (ms, n, b) in sorted(stats.items(), key=lambda kv: -kv[1][0]):
        log('  %-22s %9.3f ms/step %6d launches  %8.1f GB/s algorithmic' %
            (k, ms, n, (b / (ms * 1e-3) / 1e9) if ms > 0 else 0))
    dom = max(stats.items(), key=lambda kv: kv[1][0])[0] if stats else 'hc_ring'
    # timed region: only the dominant kernel carries HIP events
    ctx.timing(True)
    ctx.timing_only(dom)
    ctx.reset_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        isos = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.timing(False)
    ctx.timing_only(None)
    dms, dn, dbytes = ctx.kernel_stats().get(dom, (0.0, 0, 0.0))
    ms_step = (t1 - t0) * 1000.0 / args.steps
    sizes = [ctx.mesh_size(k) for k in range(3)]
    ntri = sum(s[0] for s in sizes)
    log('step %.2f ms; meshes (tris, verts): %s' % (ms_step, sizes))
    achieved = (dbytes / dn) / ((dms / dn) * 1e-3) / 1e9 if dn and dms else 0.0
    roof = {'kernel': dom, 'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
            'bytes_per_launch': dbytes / dn if dn else 0, 'avg_launch_us': dms / dn * 1e3 if dn else 0,
            'launches_per_step': dn // max(args.steps, 1),
            'traffic': traffic_from_profile(dom),
            'traffic_source': 'profiles/traffic_%s.txt (rocprofv3 --pmc, committed)' % dom}
    out = {
        'metric': 'GVoxel/s mesh extraction at 512^3 (%HBM peak)',
        'value': round(nvox / (ms_step * 1e-3) / 1e9, 4),
        'unit': 'GVoxel/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_step, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': 'large pipeline 512^3 f32: 3 Otsu isolevels, Lewiner MC33, largest cluster + '
                               'bubble fill, 100 HC iterations each', 'shape': [N, N, N],
                   'isolevels': [round(v, 4) for v in isos], 'triangles_per_step': ntri},
        'triangles_per_s': round(ntri / (ms_step * 1e-3), 1),
        'pipeline_hbm_frac': round(tot_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if stats else None,
        'kernel_ms_per_step': round(tot_ms, 3) if stats else None,
        'roofline': roof,
    }
    out['pcie_inclusive'] = pcie_inclusive(torch, vol, step, ctx)
    log('pcie-inclusive step %.2f ms (h2d %.2f, step %.2f, d2h %.2f)' % (
        out['pcie_inclusive']['ms_per_step'], out['pcie_inclusive']['h2d_ms'], out['pcie_inclusive']['step_ms'],
        out['pcie_inclusive']['d2h_ms']))
    out['host_cpu'] = host_cpu()
    atlas_vol = None
    if not args.no_atlas:
        out['atlas'], atlas_vol = atlas_leg(torch, n2m, synth)
    if not args.no_cpu:
        log('cpu baseline ...')
        out['cpu_baseline'] = cpu_baseline(vol, isos)
        log('cpu baseline %.6f GVoxel/s, stages %s' % (out['cpu_baseline']['value'], out['cpu_baseline']['stages_ms']))
        if atlas_vol is not None:
            out['atlas']['cpu_loop'] = cpu_atlas(atlas_vol)
            log('cpu atlas loop %s' % out['atlas']['cpu_loop'])
    print(json.dumps(out), flush=True)


def atlas_leg(torch, n2m, synth, reps=5):
    """Batched atlas vs the per-label GPU loop on the 256^3 / 370-label atlas."""
    t = time.perf_counter()
    vol = synth.atlas_u16()
    log('generated 256^3 atlas in %.1f s' % (time.perf_counter() - t))
    d = torch.from_numpy(vol).to('cuda')
    dims = tuple(vol.shape[::-1])
    ctx = n2m.DeviceContext()
    prm = ctx.params(0.5, cubes=n2m.CUBES_CLASSIC, preSmooth=True, onlyLargest=False, fillBubbles=False)

    def run(loop):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if ctx.atlas(d.data_ptr(), dims, prm, 0, None, loop=loop) != 0:
            raise RuntimeError('atlas failed: ' + n2m.last_error())
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0

    run(False)
    run(True)
    batched = min(run(False) for _ in range(reps))
    ent = ctx.atlas_entries()
    ntri = sum(e.nt for e in ent if e.status == n2m.ATLAS_OK)
    nok = sum(e.status == n2m.ATLAS_OK for e in ent)
    loop = min(run(True) for _ in range(2))
    log('atlas: batched %.2f ms, per-label loop %.2f ms (%d labels, %d triangles)' % (batched, loop, nok, ntri))
    ctx.timing(True)
    ctx.reset_timing()
    run(False)
    ctx.timing(False)
    for k, (ms, n, b) in sorted(ctx.kernel_stats().items(), key=lambda kv: -kv[1][0]):
        log('  atlas %-28s %8.3f ms %5d launches' % (k, ms, n))
    ctx.close()
    tuned = {k: {'tile': v[0], 'blocks_per_cu': v[1]} for k, v in n2m.atlas_tuned(0).items() if v[1] >= 0}
    log('atlas launch shapes (tuned): %s' % tuned)
    sz = np.bincount(vol.astype(np.int64).ravel())[1:]
    return {'workload': '256^3 weighted-Voronoi atlas, 370 labels (log-normal sizes: median %d, max %d voxels), '
                        'binarise 0.5 + pre-smooth + classic marching cubes' % (np.median(sz), sz.max()),
            'labels_meshed': nok, 'triangles': ntri, 'batched_ms': round(batched, 3),
            'per_label_loop_ms': round(loop, 3), 'speedup': round(loop / batched, 2),
            'batched_gvoxel_s': round(vol.size / (batched * 1e-3) / 1e9, 4), 'tuned': tuned}, vol


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic_<kernel>.txt
    (FETCH_SIZE + WRITE_SIZE per dispatch, copy-kernel calibrated), if present."""
    p = os.path.join(ROOT, 'profiles', 'traffic_%s.txt' % kernel)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith('hbm_bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


if __name__ == '__main__':
    main()
