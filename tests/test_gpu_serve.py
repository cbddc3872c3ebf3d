"""Serving form: continuous batching with HBM-budget admission.

A job's result must be the same alone and in a mix of concurrent jobs
(bit for bit: same kernels, same per-voxel summation order); jobs that arrive
while others run are admitted as memory frees up; a job larger than the pool
fails cleanly instead of blocking the queue.
"""
import threading
import time

import numpy as np
import pytest

import qtctreconstruction_amd as ct

pytestmark = pytest.mark.gpu


def _sinos(O, seed, N, nz, na):
    vol = O.phantom(seed, 8, N, N, nz)
    return np.stack([O.compute_sinogram(vol[z], na, N)[0] for z in range(nz)])


@pytest.fixture(scope="module")
def inputs(O):
    ctx = ct.Context(0)
    specs = [(48, 6, 60, ct.FilterType.Ramp), (96, 4, 90, ct.FilterType.Hann),
             (64, 9, 45, ct.FilterType.SheppLogan), (128, 3, 120, ct.FilterType.Cosine),
             (32, 17, 30, ct.FilterType.Hamming), (80, 5, 180, ct.FilterType.Bartlett)]
    jobs = []
    for k, (N, nz, na, f) in enumerate(specs):
        s = _sinos(O, 50 + k, N, nz, na)
        jobs.append(dict(N=N, nz=nz, na=na, f=f, host=s, dev=ctx.alloc(s.nbytes).upload(s),
                         ang=O.serial_angles(na)))
    yield ctx, jobs
    ctx.close()


def _run_alone(srv, j):
    t = srv.submit(j["dev"], j["ang"], j["nz"], j["na"], j["N"], j["f"])
    r = srv.wait(30000)
    assert r is not None and r[0] == t
    out = srv.download(t, (j["nz"], j["N"], j["N"]))
    srv.release(t)
    return out


def test_alone_vs_mix_bit_identical(inputs, O):
    ctx, jobs = inputs
    srv = ct.Server(pool_bytes=256 << 20, max_batch=8, slots=2)
    alone = [_run_alone(srv, j) for j in jobs]
    # the same jobs again, three times over, submitted from two threads at once
    tickets = {}
    lock = threading.Lock()

    def client(lst):
        for k in lst:
            t = srv.submit(jobs[k]["dev"], jobs[k]["ang"], jobs[k]["nz"], jobs[k]["na"],
                           jobs[k]["N"], jobs[k]["f"])
            with lock:
                tickets[t] = k

    order = list(range(len(jobs))) * 3
    th = [threading.Thread(target=client, args=(order[i::2],)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    got = 0
    while got < len(order):
        r = srv.wait(30000)
        assert r is not None, "server stalled"
        t, lat = r
        k = tickets[t]
        j = jobs[k]
        out = srv.download(t, (j["nz"], j["N"], j["N"]))
        srv.release(t)
        assert np.array_equal(out.view(np.uint32), alone[k].view(np.uint32)), k
        got += 1
    st = srv.stats()
    assert st["completed"] == len(order) + len(jobs)
    assert st["pool_used"] == 0
    srv.close()
    # and the standalone result is the reference's consumer path
    be = ct.HipBackend(0)
    for j, a in zip(jobs, alone):
        ref = be.reconstructVolumeFromSinograms(j["host"], j["ang"], ct.ReconstructionParams(j["f"], j["na"]))
        assert np.array_equal(a, ref)
        for z in range(j["nz"]):
            assert O.rel_l2(a[z], O.reconstruct_slice(j["host"][z], j["ang"], j["N"], int(j["f"]))) <= 1e-5


def test_split_bf16_filter_backend(inputs, O):
    """set_filter_backend(GEMM_SPLIT_BF16): each job's result equals a
    standalone reconstructVolumeFromSinograms on a context with the same
    backend bit for bit, and stays within 1e-4 of the FP64 oracle; the
    default FFT backend comes back after the switch back."""
    ctx, jobs = inputs
    srv = ct.Server(pool_bytes=256 << 20, max_batch=8, slots=2)
    be = ct.HipBackend(0)
    be.ctx.set_filter_backend(ct.FilterBackend.GEMM_SPLIT_BF16)
    srv.set_filter_backend(ct.FilterBackend.GEMM_SPLIT_BF16)
    tickets = [srv.submit(j["dev"], j["ang"], j["nz"], j["na"], j["N"], j["f"]) for j in jobs]
    done = {}
    while len(done) < len(tickets):
        r = srv.wait(30000)
        assert r is not None
        done[r[0]] = True
    for t, j in zip(tickets, jobs):
        got = srv.download(t, (j["nz"], j["N"], j["N"]))
        srv.release(t)
        ref = be.reconstructVolumeFromSinograms(j["host"], j["ang"],
                                                ct.ReconstructionParams(j["f"], j["na"]))
        assert np.array_equal(got, ref), j["N"]
        ref64 = O.reconstruct_slice64(j["host"][0], j["ang"], j["N"], int(j["f"]))
        assert O.rel_l2(got[0], ref64) <= 1e-4, j["N"]
    srv.set_filter_backend(ct.FilterBackend.FFT_F32)
    fft = _run_alone(srv, jobs[0])
    be.ctx.set_filter_backend(ct.FilterBackend.FFT_F32)
    ref = be.reconstructVolumeFromSinograms(jobs[0]["host"], jobs[0]["ang"],
                                            ct.ReconstructionParams(jobs[0]["f"], jobs[0]["na"]))
    assert np.array_equal(fft, ref)
    with pytest.raises(ct.BackendError):
        srv.set_filter_backend(9)
    srv.close()


def test_admission_by_budget_and_overlap(inputs):
    ctx, jobs = inputs
    j = jobs[3]  # 128^2 x 3 slices x 120 angles
    out_b = 4 * j["nz"] * j["N"] * j["N"]
    filt_b = 16 * 1 * j["na"] * int(np.ceil(np.sqrt(2) * j["N"]))
    pool = 3 * (out_b + filt_b) + 4096  # room for three jobs at a time
    srv = ct.Server(pool_bytes=pool, max_batch=4, slots=2)
    n = 12
    ts = [srv.submit(j["dev"], j["ang"], j["nz"], j["na"], j["N"], j["f"]) for _ in range(n)]
    seen = 0
    while seen < n:
        r = srv.wait(30000)
        assert r is not None, "admission stalled"
        srv.release(r[0])
        seen += 1
    st = srv.stats()
    assert st["max_inflight_jobs"] <= 3
    assert st["completed"] == n
    srv.close()


def test_oversize_job_fails_cleanly(inputs):
    ctx, jobs = inputs
    j = jobs[3]  # needs ~196 KB of output + ~349 KB of filtered sinogram
    srv = ct.Server(pool_bytes=256 << 10, max_batch=4, slots=1)
    t = srv.submit(j["dev"], j["ang"], j["nz"], j["na"], j["N"], j["f"])
    with pytest.raises(ct.BackendError):
        srv.wait(30000)
    small = jobs[4]  # 32^2: fits after the big one was refused
    srv2 = ct.Server(pool_bytes=64 << 20, max_batch=4, slots=1)
    t2 = srv2.submit(small["dev"], small["ang"], small["nz"], small["na"], small["N"], small["f"])
    assert srv2.wait(30000)[0] == t2
    srv.close()
    srv2.close()


def test_policies_give_identical_results(inputs):
    ctx, jobs = inputs
    outs = {}
    for policy, cap in ((ct.SERVE_FIFO, 0.0), (ct.SERVE_SIZE_AGED, 0.05), (ct.SERVE_SIZE_AGED, 0.0)):
        srv = ct.Server(pool_bytes=256 << 20, max_batch=8, slots=2)
        srv.set_policy(policy, batch_cap_ms=cap)
        tick = {srv.submit(j["dev"], j["ang"], j["nz"], j["na"], j["N"], j["f"]): k
                for k, j in enumerate(jobs)}
        for _ in jobs:
            r = srv.wait(30000)
            assert r is not None, "server stalled"
            k = tick[r[0]]
            j = jobs[k]
            out = srv.download(r[0], (j["nz"], j["N"], j["N"]))
            srv.release(r[0])
            if k in outs:
                assert np.array_equal(out.view(np.uint32), outs[k].view(np.uint32)), (policy, k)
            else:
                outs[k] = out
        srv.close()


def test_set_policy_rejects_bad_arguments():
    srv = ct.Server(pool_bytes=1 << 20, max_batch=2, slots=1)
    for args in ((7, 5.0, 0.25), (ct.SERVE_FIFO, -1.0, 0.25), (ct.SERVE_SIZE_AGED, 5.0, -0.5)):
        with pytest.raises(ct.BackendError):
            srv.set_policy(*args)
    srv.close()


def test_size_order_lets_small_jobs_pass_a_large_one():
    """One batch per job, one slot: while a long job runs, a second long job
    and three small ones queue. FIFO runs them in submission order; the
    size-ordered policy runs the small ones before the queued long one."""
    ctx = ct.Context(0)
    N, nz, na = 768, 512, 720  # ~2e11 voxel-updates: tens of ms on the device
    big = ctx.alloc(4 * nz * na * N).upload(np.zeros((nz, na, N), np.float32))
    n, snz, sna = 64, 4, 60
    small = ctx.alloc(4 * snz * sna * n).upload(np.ones((snz, sna, n), np.float32))
    ang_big = (np.float32(180.0 / na) * np.arange(na, dtype=np.float32)).astype(np.float32)
    ang_small = (np.float32(180.0 / sna) * np.arange(sna, dtype=np.float32)).astype(np.float32)
    orders = {}
    for policy in (ct.SERVE_FIFO, ct.SERVE_SIZE_AGED):
        srv = ct.Server(pool_bytes=8 << 30, max_batch=1, slots=1)
        srv.set_policy(policy, batch_cap_ms=5.0, age_weight=0.25)
        name = {srv.submit(big, ang_big, nz, na, N, 0): "B1"}
        t0 = time.time()
        while srv.stats()["batches"] < 1:  # B1 is admitted alone and runs
            assert time.time() - t0 < 30, "B1 not admitted"
            time.sleep(0.0005)
        name[srv.submit(big, ang_big, nz, na, N, 0)] = "B2"
        for k in range(3):
            name[srv.submit(small, ang_small, snz, sna, n, 0)] = f"S{k}"
        seq = []
        for _ in range(5):
            r = srv.wait(60000)
            assert r is not None, "server stalled"
            seq.append(name[r[0]])
            srv.release(r[0])
        srv.close()
        orders[policy] = seq
    assert orders[ct.SERVE_FIFO] == ["B1", "B2", "S0", "S1", "S2"]
    assert orders[ct.SERVE_SIZE_AGED] == ["B1", "S0", "S1", "S2", "B2"]
    big.free()
    small.free()
    ctx.close()
