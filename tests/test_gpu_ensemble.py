"""GPU tests of the batched ensemble and the served ensemble.

* every cluster's result in a batch is bitwise identical to the same cluster
  run alone (its decomposition depends only on its own N and target count;
  under checked precision a cluster's flagged targets are recomputed once
  all its target blocks are done, with a plan for their total count), and
  meets the oracle bars of test_gpu_parity;
* the captured graph replays correctly with eager work between replays and a
  different mix each time;
* the server completes concurrently submitted requests with the same bits,
  refuses admission beyond its HBM budget, and reports latency percentiles;
* the batched-ensemble workload (256 clusters) and a served mix with
  clusters up to 1048576 particles at their sizes.
"""
import sys
import threading

import numpy as np
import pytest

import oracle
from oracle.compare import compare_lists, force_errors
from abyss_amd import ic, regf

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_clusters(sizes, seed0=100, k_nb=32):
    out = []
    for k, n in enumerate(sizes):
        rng = np.random.default_rng(seed0 + k)
        if k % 3 == 2:
            x, v, m = ic.king(n, W0=3.0 + 6.0 * rng.uniform(), seed=seed0 + k,
                              masses=ic.kroupa_masses(n, rng))
        else:
            x, v, m = ic.plummer(n, seed=seed0 + k, masses=ic.kroupa_masses(n, rng))
        r2 = ic.neighbour_radius2(x, min(k_nb, n - 2))
        out.append((x, v, m, r2))
    return out


def random_targets(n, seed):
    rng = np.random.default_rng(seed)
    size = int(np.exp(rng.uniform(0.0, np.log(n + 1))))
    size = max(1, min(n, size))
    return rng.permutation(n)[:size].astype(np.int32)


SIZES = [300, 1024, 2000, 257, 5000, 9000, 20000, 700, 3000, 12000, 65536, 1500]


@pytest.fixture(scope="module")
def clusters():
    return make_clusters(SIZES)


def check_vs_oracle(cl, t, res, label):
    x, v, m, r2 = cl
    acc, adot, nnb, lst = res
    ra, rj, rn, rl, sa, sj = oracle.regf(x, v, m, r2, t, lst.shape[1], scales=True)
    err = force_errors(acc, adot, ra, rj, sa, sj)
    lists = compare_lists(nnb, lst, rn, rl, x, r2, t)
    assert lists["hard_mismatches"] == 0 and lists["not_ascending"] == 0, (label, lists)
    assert err["a_norm"]["max"] <= 1e-6 and err["jdot_norm"]["max"] <= 1e-6, (label, err)
    assert err["a_rel"]["p99"] <= 1e-6, (label, err)


def test_ensemble_matches_alone_bitwise(clusters):
    say("case: ensemble mix vs alone")
    with regf.Ensemble(device=0, cap=128, max_clusters=len(clusters)) as ens:
        for c, (x, v, m, r2) in enumerate(clusters):
            ens.upload(c, m, x, v, r2)
        targets = [random_targets(len(cl[2]), 7 + c) for c, cl in enumerate(clusters)]
        for c, t in enumerate(targets):
            ens.set_targets(c, t)
        ens.launch()
        ens.sync()
        mixed = [ens.fetch(c) for c in range(len(clusters))]
        for c, t in enumerate(targets):
            for k in range(len(clusters)):
                ens.set_targets(k, t if k == c else [])
            ens.launch()
            ens.sync()
            alone = ens.fetch(c)
            for a, b in zip(alone, mixed[c]):
                assert np.array_equal(a, b), f"cluster {c} differs alone vs mixed"
            check_vs_oracle(clusters[c], t, mixed[c], f"cluster{c}")
            say(f"  cluster {c}: n={len(clusters[c][2])} nt={len(t)} identical alone/mixed")


def test_ensemble_checked_precision(clusters):
    """Checked precision in a batch: lists and counts those of split
    precision, every target within 1e-6 relative of the oracle (a and jdot)
    on rows whose regular set the FP64 test agrees with, recomputed targets
    FP64-accurate, a cluster alone equal to the mix bit for bit, and split
    precision restored bit for bit after switching back."""
    say("case: ensemble checked precision")
    with regf.Ensemble(device=0, cap=128, max_clusters=len(clusters)) as ens:
        for c, (x, v, m, r2) in enumerate(clusters):
            ens.upload(c, m, x, v, r2)
        targets = [random_targets(len(cl[2]), 51 + c) for c, cl in enumerate(clusters)]
        targets[10] = np.arange(len(clusters[10][2]), dtype=np.int32)   # one full cluster
        for c, t in enumerate(targets):
            ens.set_targets(c, t)
        ens.set_precision(regf.PREC_SPLIT)
        ens.launch()
        ens.sync()
        split = [ens.fetch(c) for c in range(len(clusters))]
        ens.set_precision(regf.PREC_CHECKED)
        ens.launch()
        ens.sync()
        chk = [ens.fetch(c) for c in range(len(clusters))]
        nre = [ens.last_recheck(c) for c in range(len(clusters))]
        # alone vs mixed for the full cluster and a small one
        for c in (10, 3):
            for k in range(len(clusters)):
                ens.set_targets(k, targets[k] if k == c else [])
            ens.launch()
            ens.sync()
            alone = ens.fetch(c)
            assert ens.last_recheck(c) == nre[c]
            for a, b in zip(alone, chk[c]):
                assert np.array_equal(a, b), f"cluster {c} differs alone vs mixed (checked)"
        ens.set_precision(regf.PREC_SPLIT)
        for c, t in enumerate(targets):
            ens.set_targets(c, t)
        ens.launch()
        ens.sync()
        again = [ens.fetch(c) for c in range(len(clusters))]
    say(f"  recomputed per cluster: {nre}")
    assert sum(nre) > 0
    for c, (cl, t) in enumerate(zip(clusters, targets)):
        for a, b in zip(again[c], split[c]):
            assert np.array_equal(a, b), f"cluster {c}: split precision not restored"
        assert np.array_equal(chk[c][2], split[c][2]) and np.array_equal(chk[c][3], split[c][3])
        changed = np.any(chk[c][0] != split[c][0], axis=1) | np.any(chk[c][1] != split[c][1],
                                                                     axis=1)
        assert changed.sum() <= nre[c]
        x, v, m, r2 = cl
        ra, rj, rn, rl = oracle.regf(x, v, m, r2, t, 128)
        w = 128
        valid = np.arange(w)[None, :] < np.minimum(np.abs(chk[c][2]), w)[:, None]
        rows = (chk[c][2] == rn) & np.all((chk[c][3] == rl) | ~valid, axis=1)
        err = force_errors(chk[c][0][rows], chk[c][1][rows], ra[rows], rj[rows])
        assert err["a_rel"]["max"] <= 1e-6 and err["jdot_rel"]["max"] <= 1e-6, (c, err)
        if (changed & rows).any():
            rerr = force_errors(chk[c][0][changed & rows], chk[c][1][changed & rows],
                                ra[changed & rows], rj[changed & rows])
            assert rerr["a_rel"]["max"] <= 1e-11 and rerr["jdot_rel"]["max"] <= 1e-11, (c, rerr)
        say(f"  cluster {c}: n={len(m)} nt={len(t)} recomputed={nre[c]} band_rows="
            f"{(~rows).sum()} jdot_rel_max={err['jdot_rel']['max']:.2e}")


def test_ensemble_graph_replays_with_changing_mix(clusters):
    say("case: ensemble graph replays")
    rng = np.random.default_rng(3)
    with regf.Ensemble(device=0, cap=128, max_clusters=len(clusters)) as ens:
        for c, (x, v, m, r2) in enumerate(clusters):
            ens.upload(c, m, x, v, r2)
        for rep in range(4):
            chosen = sorted(rng.permutation(len(clusters))[:5].tolist())
            tg = {}
            for c in range(len(clusters)):
                if c in chosen:
                    tg[c] = random_targets(len(clusters[c][2]), 1000 * rep + c)
                    ens.set_targets(c, tg[c])
                else:
                    ens.set_targets(c, [])
            ens.launch()
            ens.sync()
            # eager single-cluster work between replays
            with regf.Context(device=0, cap=128) as ctx:
                x, v, m, r2 = clusters[0]
                ctx.upload(m, x, v, r2)
                ctx.query(np.arange(64, dtype=np.int32))
            for c in chosen:
                check_vs_oracle(clusters[c], tg[c], ens.fetch(c), f"rep{rep}/c{c}")
            say(f"  replay {rep}: clusters {chosen} ok")


@pytest.mark.parametrize("prec", [regf.PREC_CHECKED, regf.PREC_SPLIT])
def test_ensemble_block_ranges_give_same_bits(clusters, prec):
    """A cluster's target blocks launched over several batches (in any
    order) give the bits of one launch; under checked precision the flagged
    targets of all batches are recomputed together once the last block is
    done."""
    say(f"case: ensemble block ranges, precision {prec}")
    x, v, m, r2 = clusters[10]
    t = np.random.default_rng(11).permutation(len(m))[:20000].astype(np.int32)
    with regf.Ensemble(device=0, cap=128, max_clusters=2) as ens:
        ens.set_precision(prec)
        ens.upload(0, m, x, v, r2)
        ens.set_targets(0, t)
        ens.launch()
        ens.sync()
        whole = ens.fetch(0)
        nre = ens.last_recheck(0)
        for ranges in (((0, 3), (3, None), (-1, -1)), ((-1, -1), (0, 3), (3, None))):
            ens.set_targets(0, t)
            nb, tpb = ens.blocks(0)
            say(f"  {nb} blocks of {tpb} targets")
            for first, count in ranges:
                if first < 0:
                    first, count = 3 + nb // 2, -1
                elif count is None:
                    count = nb // 2
                ens.set_block_range(0, first, count)
                ens.launch()
                ens.sync()
            sliced = ens.fetch(0)
            assert ens.last_recheck(0) == nre
            for a, b in zip(whole, sliced):
                assert np.array_equal(a, b)
    if prec == regf.PREC_CHECKED:
        assert nre > 0


@pytest.mark.parametrize("budget", [3e10, 2e7])
def test_server_concurrent_requests_match_alone(clusters, budget):
    say(f"case: server concurrent requests, batch budget {budget:g} pairs")
    sub = clusters[:8]
    with regf.Server(device=0, cap=128, max_clusters=8, hbm_budget=8 << 30) as srv:
        srv.set_batch_budget(budget)
        ids = [srv.add_cluster(m, x, v, r2) for (x, v, m, r2) in sub]
        reqs = [(ids[c % 8], random_targets(len(sub[c % 8][2]), 500 + c)) for c in range(40)]
        results = [None] * len(reqs)

        def client(lo, hi):
            for q in range(lo, hi):
                tk, nt = srv.submit(*reqs[q])
                results[q] = srv.wait(tk, nt)

        th = [threading.Thread(target=client, args=(k * 10, k * 10 + 10)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
        say(f"  stats {st}")
        assert st["completed"] == len(reqs)
        assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
    with regf.Ensemble(device=0, cap=128, max_clusters=8) as ens:
        for c, (x, v, m, r2) in enumerate(sub):
            ens.upload(c, m, x, v, r2)
        for q, (cid, t) in enumerate(reqs):
            c = ids.index(cid)
            for k in range(8):
                ens.set_targets(k, t if k == c else [])
            ens.launch()
            ens.sync()
            alone = ens.fetch(c)
            got = results[q]
            for a, b in zip(alone[:3], got[:3]):
                assert np.array_equal(a, b), f"request {q} differs from the cluster alone"
            for k in range(len(t)):   # list entries past the count are unspecified
                nk = alone[2][k]
                assert np.array_equal(alone[3][k, :nk], got[3][k, :nk]), f"request {q} list"


def test_server_checked_precision_matches_alone(clusters):
    """Served requests under checked precision, with a small batch budget so
    requests span several batches: every result bit for bit that of the
    cluster alone (checked ensemble), since the flagged targets of a request
    are recomputed together after its last block."""
    say("case: server checked precision vs alone")
    sub = clusters[:8]
    with regf.Server(device=0, cap=128, max_clusters=8, hbm_budget=8 << 30) as srv:
        srv.set_precision(regf.PREC_CHECKED)
        srv.set_batch_budget(2e7)
        ids = [srv.add_cluster(m, x, v, r2) for (x, v, m, r2) in sub]
        reqs = [(ids[c % 8], random_targets(len(sub[c % 8][2]), 900 + c)) for c in range(16)]
        tickets = [srv.submit(*r) for r in reqs]
        results = [srv.wait(tk, nt) for tk, nt in tickets]
        assert srv.stats()["batches"] > len(set(ids))
    differ = 0
    with regf.Ensemble(device=0, cap=128, max_clusters=8) as ens:
        ens.set_precision(regf.PREC_CHECKED)
        for c, (x, v, m, r2) in enumerate(sub):
            ens.upload(c, m, x, v, r2)
        for q, (cid, t) in enumerate(reqs):
            c = ids.index(cid)
            for k in range(8):
                ens.set_targets(k, t if k == c else [])
            ens.launch()
            ens.sync()
            alone = ens.fetch(c)
            got = results[q]
            assert np.array_equal(alone[2], got[2]), f"request {q} counts"
            for k in range(len(t)):
                nk = alone[2][k]
                assert np.array_equal(alone[3][k, :nk], got[3][k, :nk]), f"request {q} list"
            for a, b in zip(alone[:2], got[:2]):
                differ += int((~np.all(a == b, axis=1)).sum())
    say(f"  targets differing from the cluster alone: {differ}")
    assert differ == 0


def test_server_admission_by_budget(clusters):
    say("case: server admission")
    x, v, m, r2 = clusters[6]             # n = 20000
    probe = regf.Server(device=0, cap=128, max_clusters=4, hbm_budget=1 << 40)
    need = probe.footprint(len(m))
    probe.close()
    with regf.Server(device=0, cap=128, max_clusters=4, hbm_budget=int(need * 1.5)) as srv:
        a = srv.add_cluster(m, x, v, r2)
        with pytest.raises(regf.AbyssError) as e:
            srv.add_cluster(m, x, v, r2)
        assert e.value.code == regf.ERR_ADMISSION
        srv.remove_cluster(a)
        b = srv.add_cluster(m, x, v, r2)
        tk, nt = srv.submit(b, np.arange(10, dtype=np.int32))
        res = srv.wait(tk, nt)
        check_vs_oracle(clusters[6], np.arange(10, dtype=np.int32), res[:4], "admitted")


def test_server_request_admission(clusters):
    say("case: server request admission (outstanding-work limit)")
    x, v, m, r2 = clusters[10]             # n = 65536
    n = len(m)
    big = np.arange(n, dtype=np.int32)     # 65536 x 65535 pairs
    small = np.arange(16, dtype=np.int32)
    with regf.Server(device=0, cap=128, max_clusters=2, hbm_budget=8 << 30) as srv:
        c = srv.add_cluster(m, x, v, r2)
        srv.set_batch_budget(2e7)          # the big request spans ~200 batches (ms)
        srv.set_admission(1.0)             # only a lone request is admitted
        tk_big, nt_big = srv.submit(c, big)
        assert tk_big is not None          # nothing outstanding: admitted
        tk2, _ = srv.submit(c, small)      # microseconds later the big one still runs
        srv.wait(tk_big, nt_big, copy=False)
        st = srv.stats()
        say(f"  stats {st}")
        assert tk2 is None and st["rejected"] == 1
        assert st["peak_outstanding_pairs"] == float(n) * (n - 1)
        # the queue has drained: the next request is admitted and correct
        tk3, nt3 = srv.submit(c, small)
        assert tk3 is not None
        check_vs_oracle(clusters[10], small, srv.wait(tk3, nt3)[:4], "admitted after drain")
        srv.set_admission(0)               # unlimited again: a burst all queues
        tks = [srv.submit(c, small) for _ in range(5)]
        assert all(t is not None for t, _ in tks)
        for t, nt in tks:
            srv.wait(t, nt, copy=False)
        st = srv.stats()
        assert st["rejected"] == 1 and st["completed"] == 7 and st["max_pending"] >= 1
        srv.reset_stats()
        assert srv.stats()["rejected"] == 0


def _sample_vs_oracle(cl, t, res, label, nsample=8, seed=0):
    """a and jdot of `nsample` targets within 1e-6 relative of the oracle
    (the default precision's contract); lists exact outside the band."""
    x, v, m, r2 = cl[:4]
    acc, adot, nnb, lst = res[:4]
    cap = lst.shape[1]
    pick = np.random.default_rng(seed).permutation(len(t))[:nsample]
    ts = t[pick]
    ra, rj, rn, rl = oracle.regf(x, v, m, r2, ts, cap)
    lists = compare_lists(nnb[pick], lst[pick], rn, rl, x, r2, ts)
    w = cap
    valid = np.arange(w)[None, :] < np.minimum(np.abs(nnb[pick]), w)[:, None]
    rows = (nnb[pick] == rn) & np.all((lst[pick] == rl) | ~valid, axis=1)
    err = force_errors(acc[pick][rows], adot[pick][rows], ra[rows], rj[rows])
    assert lists["hard_mismatches"] == 0 and lists["not_ascending"] == 0, (label, lists)
    assert err["a_rel"]["max"] <= 1e-6 and err["jdot_rel"]["max"] <= 1e-6, (label, err)
    return err


def test_config4_ensemble_256_clusters():
    """The batched-ensemble workload at its size: 256 Plummer clusters, N
    log-uniform in [1024, 65536], active fraction in [0.01, 1], one batched
    launch (default precision).  Every cluster's counts and lists equal
    those of its own planned launch, a and jdot of both agree to the
    contract, 16 sampled targets per cluster meet 1e-6 against the oracle,
    and eight clusters run alone give the bits of the mix."""
    from abyss_amd import workloads
    cl = workloads.ensemble_clusters()
    say(f"case: config 4, {len(cl)} clusters, sum N {sum(len(c[2]) for c in cl)}")
    cap = 256
    with regf.Ensemble(device=0, cap=cap, max_clusters=len(cl)) as ens:
        for k, (x, v, m, r2, t) in enumerate(cl):
            ens.upload(k, m, x, v, r2)
            ens.set_targets(k, t)
        ens.launch()
        ens.sync()
        mixed = [ens.fetch(k) for k in range(len(cl))]
        nre = sum(ens.last_recheck(k) for k in range(len(cl)))
        order = np.argsort([len(c[2]) for c in cl])
        alone_set = sorted(set([int(order[0]), int(order[-1])] +
                               np.random.default_rng(4).permutation(len(cl))[:6].tolist()))
        for k in alone_set:
            for j in range(len(cl)):
                ens.set_targets(j, cl[j][4] if j == k else [])
            ens.launch()
            ens.sync()
            alone = ens.fetch(k)
            for a, b in zip(alone, mixed[k]):
                assert np.array_equal(a, b), f"cluster {k} differs alone vs mixed"
    say(f"  recomputed {nre}; alone == mixed for clusters {alone_set}")
    worst = 0.0
    with regf.Context(device=0, cap=cap) as c:
        for k, (x, v, m, r2, t) in enumerate(cl):
            c.upload(m, x, v, r2)
            own = c.query(t)
            assert np.array_equal(own[2], mixed[k][2]), f"cluster {k} counts"
            w = np.arange(cap)[None, :] < np.minimum(np.abs(own[2]), cap)[:, None]
            assert np.all((own[3] == mixed[k][3]) | ~w), f"cluster {k} lists"
            d = force_errors(mixed[k][0], mixed[k][1], own[0], own[1])
            worst = max(worst, d["a_rel"]["max"], d["jdot_rel"]["max"])
            _sample_vs_oracle((x, v, m, r2), t, mixed[k], f"config4/{k}", nsample=16, seed=k)
    say(f"  batch vs planned launches: max rel diff {worst:.2e}")
    assert worst <= 2e-6


def test_config5_served_mix_with_large_clusters():
    """A served mix at the served workload's sizes: 16 clusters, N
    log-uniform in [4096, 1048576] plus a King cluster of 524288 and a
    Plummer cluster of 1048576 particles; 48 regular-step requests (active
    blocks log-uniform in [1, min(N, 32768)]) submitted by four threads at
    Poisson times, under the default precision.  Every request's results
    equal the same request run alone on the server bit for bit, equal plain
    batched-ensemble launches (other work items) to 1e-12 on 99.9 % of the
    targets and within the contract on all (lists exactly), and meet the
    contract against the oracle on sampled targets."""
    import time
    from abyss_amd import workloads
    sizes = list(np.exp(np.random.default_rng(5).uniform(np.log(4096), np.log(1048576),
                                                         16)).astype(int))
    sizes[1] = 524288        # odd index: a King cluster
    sizes[2] = 1048576       # even index: a Plummer cluster
    cl = workloads.served_clusters(sizes=sizes, seed=5)
    say(f"case: config 5 served mix, sizes {sizes}")
    cap = 128
    rng = np.random.default_rng(17)
    reqs = []
    for q in range(48):
        c = int(rng.integers(len(cl)))
        if q < 4:
            c = 1 + (q % 2)   # the two large clusters are always in the mix
        n = len(cl[c][2])
        size = int(min(n, 32768, max(1, np.exp(rng.uniform(0.0, np.log(min(n, 32768) + 1))))))
        reqs.append((c, rng.choice(n, size, replace=False).astype(np.int32)))
    results = [None] * len(reqs)
    with regf.Server(device=0, cap=cap, max_clusters=len(cl), hbm_budget=64 << 30) as srv:
        ids = [srv.add_cluster(m, x, v, r2) for (x, v, m, r2) in cl]

        def client(part, seed):
            r = np.random.default_rng(seed)
            tickets = []
            for q in part:
                time.sleep(r.exponential(1.0 / 100.0))
                tk, nt = srv.submit(ids[reqs[q][0]], reqs[q][1])
                tickets.append((q, tk, nt))
            for q, tk, nt in tickets:
                results[q] = srv.wait(tk, nt)

        th = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)), 40 + k))
              for k in range(4)]
        for t_ in th:
            t_.start()
        for t_ in th:
            t_.join()
        st = srv.stats()
        say(f"  stats {st}")
        assert st["completed"] == len(reqs)
        # each request alone on the same server: the same bits
        for q, (c, t) in enumerate(reqs):
            tk, nt = srv.submit(ids[c], t)
            alone = srv.wait(tk, nt)
            got = results[q]
            for a, b in zip(alone[:3], got[:3]):
                assert np.array_equal(a, b), f"request {q} (cluster {c}) differs from alone"
            for k in range(len(t)):
                nk = min(abs(int(alone[2][k])), cap)
                assert np.array_equal(alone[3][k, :nk], got[3][k, :nk]), f"request {q} list"
    # and as plain batched-ensemble launches (longer work items, so other
    # j-splits): the same lists, sums within 1e-12
    with regf.Ensemble(device=0, cap=cap, max_clusters=len(cl)) as ens:
        for c, (x, v, m, r2) in enumerate(cl):
            ens.upload(c, m, x, v, r2)
        for q, (c, t) in enumerate(reqs):
            for k in range(len(cl)):
                ens.set_targets(k, t if k == c else [])
            ens.launch()
            ens.sync()
            alone = ens.fetch(c)
            got = results[q]
            assert np.array_equal(alone[2], got[2]), f"request {q} counts"
            for a, b in zip(alone[:2], got[:2]):
                # other FP64 summation orders; a target at the checked
                # precision's flag threshold may be recomputed in one and
                # not the other, so a few may differ within the contract
                na = np.linalg.norm(a, axis=1)
                rel = np.linalg.norm(b - a, axis=1) / np.where(na > 0, na, 1.0)
                assert rel.max() <= 2e-6 and np.mean(rel <= 1e-12) >= 0.999, (q, rel.max())
            for k in range(len(t)):
                nk = min(abs(int(alone[2][k])), cap)
                assert np.array_equal(alone[3][k, :nk], got[3][k, :nk]), f"request {q} list"
            _sample_vs_oracle(cl[c], t, got, f"config5/req{q}", nsample=8, seed=q)
    say("  every request equals itself alone and meets the contract")
