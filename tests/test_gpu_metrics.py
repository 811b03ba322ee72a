"""Structure scores on the device (cg_alignment_counts) against the dense numpy
restatement of benchmark/experiment.py:115-141, and the experiment harness
end to end."""

import json

import networkx as nx
import numpy as np
import pytest
import torch

from cugal_amd import Config, HungarianMethod, SinkhornMethod
from cugal_amd.experiment import (Algorithm, Experiment, ExperimentResults, Graph, GraphKind, NoiseLevel,
                                  generate_graph, structure_scores)
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _oracle_scores(source, target, answer, mapping):
    S = nx.to_numpy_array(source) if isinstance(source, nx.Graph) else source
    T = nx.to_numpy_array(target) if isinstance(target, nx.Graph) else target
    return (oracle.induced_conserved_structure(S, T, answer, mapping), oracle.edge_correctness(S, T, answer, mapping),
            oracle.symmetric_substructure(S, T, answer, mapping))


def _instance(n, p, noise, seed):
    # Graphs as the harness builds them: from edge lists, so their node order
    # is the order of first appearance, not the labels' order.
    gen = np.random.default_rng(seed)
    g = nx.erdos_renyi_graph(n, p, seed=seed)
    source, target, mapping, _ = generate_graph(g, None, gen, NoiseLevel(0.0, noise, True))
    m = len(mapping)
    source.add_nodes_from(range(m))
    target.add_nodes_from(range(m))
    return source, target, mapping


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_scores_match_oracle(device, seed):
    source, target, mapping = _instance(300, 0.03, 0.05, seed)
    assert list(target.nodes()) != sorted(target.nodes())
    rng = np.random.default_rng(seed)
    m = len(mapping)
    answers = {
        "random": rng.permutation(m),
        "truth": np.asarray(mapping),
        "identity": np.arange(m),
        "duplicates": np.where(rng.random(m) < 0.2, 0, rng.permutation(m)),
        "negative": np.where(rng.random(m) < 0.1, -1 - rng.integers(0, m, m), np.arange(m)),
    }
    for name, ans in answers.items():
        got = structure_scores(source, target, ans, mapping, device)
        want = _oracle_scores(source, target, ans, mapping)
        assert got == tuple(float(x) for x in want), (name, got, want)


def test_scores_errors_and_special_graphs(device):
    source, target, mapping = _instance(120, 0.05, 0.0, 4)
    m = len(mapping)
    # Out-of-range answers and length mismatches score 0, as the reference's
    # try/except does.
    bad = np.arange(m)
    bad[3] = m
    assert structure_scores(source, target, bad, mapping, device) == (0.0, 0.0, 0.0)
    assert _oracle_scores(source, target, bad, mapping) == (0, 0, 0)
    assert structure_scores(source, target, np.arange(m - 1), mapping, device) == (0.0, 0.0, 0.0)
    # Self-loops are one entry of the dense matrix; a directed pair; dense input.
    source.add_edge(5, 5)
    target.add_edge(7, 7)
    ans = np.random.default_rng(1).permutation(m)
    assert structure_scores(source, target, ans, mapping, device) == _oracle_scores(source, target, ans, mapping)
    ds = nx.DiGraph(nx.gnp_random_graph(60, 0.1, seed=3, directed=True))
    dt = nx.DiGraph(nx.gnp_random_graph(60, 0.1, seed=4, directed=True))
    perm = np.random.default_rng(2).permutation(60)
    ans = np.random.default_rng(3).permutation(60)
    assert structure_scores(ds, dt, ans, perm, device) == _oracle_scores(ds, dt, ans, perm)
    S, T = nx.to_numpy_array(source), nx.to_numpy_array(target)
    ans = np.random.default_rng(5).permutation(m)
    assert structure_scores(S, T, ans, mapping, device) == _oracle_scores(S, T, ans, mapping)
    with pytest.raises(ValueError):
        structure_scores(source, target, ans, np.zeros(m, dtype=np.int64), device)


def test_standard_scores_of_the_true_alignment(device):
    source, target, mapping = _instance(200, 0.04, 0.0, 6)
    # Noise free: the true alignment conserves every edge.
    assert structure_scores(source, target, mapping, mapping, device, standard=True) == (1.0, 1.0, 1.0)
    assert structure_scores(source, target, np.arange(len(mapping)), mapping, device, standard=False) != (1.0,) * 3


def test_experiment_runs_end_to_end(device, tmp_path):
    # Two predefined edge-list files and a Newman-Watts family, two noise
    # levels, two algorithms; results written and read back as JSON.
    g = nx.watts_strogatz_graph(120, 6, 0.1, seed=1)
    src_file, tgt_file = tmp_path / "s.txt", tmp_path / "t.txt"
    src_file.write_text("# source\n" + "\n".join(f"{u} {v}" for u, v in g.edges()) + "\n")
    # The reference takes a predefined pair as aligned by the identity and
    # permutes the target itself (benchmark/experiment.py:313-324).
    tgt_file.write_text("\n".join(f"{u}\t{v}" for u, v in g.edges()) + "\n")
    base = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                  sinkhorn_regularization=0.3, mu=2.0, iter_count=4, frank_wolfe_iter_count=4,
                  use_sparse_adjacency=True)
    algos = [Algorithm(base, False),
             Algorithm(Config(**{**base.__dict__, "hungarian_method": HungarianMethod.SCIPY}), False)]
    exp = Experiment(algorithms=algos,
                     graphs=[Graph(GraphKind.NEWMAN_WATTS, {"n": 150, "k": 6, "p": 0.1}),
                             Graph(GraphKind.PREDEFINED_GRAPHS, {"source_file": str(src_file),
                                                                 "target_file": str(tgt_file)})],
                     noise_levels=[NoiseLevel(0.0, 0.0, True), NoiseLevel(0.0, 0.05, True)], seed=11, num_runs=1)
    res = exp.run()
    assert len(res.results) == 2 and all(len(r) == 2 and all(len(x) == 2 for x in r) for r in res.results)
    for graph, level, algo, result in res.all_results():
        for v in (result.ics, result.ec, result.sss, result.accuracy):
            assert 0.0 <= v <= 1.0
        assert result.profile.frank_wolfe_iterations == 16
    res.dump(str(tmp_path))
    back = ExperimentResults.from_dict(json.load(open(tmp_path / (res.name() + ".json"))))
    assert back.to_dict() == res.to_dict()


def _published_config(device_cfg: bool, **kw):
    """The reference's published GPU run (CUGAL_res.json: LOG, reg 0.3, <= 500
    iterations, thr 1e-3, mu 2, 15 x 10 FW, sparse, SCIPY)."""
    base = dict(sinkhorn_regularization=0.3, sinkhorn_threshold=1e-3, mu=2.0, iter_count=15,
                frank_wolfe_iter_count=10, use_sparse_adjacency=True)
    base.update(kw)
    if device_cfg:
        return Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                      sinkhorn_iterations=500, hungarian_method=HungarianMethod.SCIPY, **base)
    return oracle.OracleConfig(sinkhorn_method="LOG", sinkhorn_iterations=500, hungarian_method="SCIPY", **base)


def _oracle_run(source, target, cfg, threads=8):
    # a fixed thread count: torch's CPU reductions split by thread count
    import copy

    before = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        return oracle.align(copy.deepcopy(source), copy.deepcopy(target), cfg)
    finally:
        torch.set_num_threads(before)


def _first_count_difference(device_profiles, run, states, n, reg, threshold, eval_freq=10):
    """The first solve whose iteration count differs between the device and
    the oracle, checked to be a threshold crossing within FP32 rounding of the
    device's own state.  Equal counts do not mean equal states (solves that
    run to the iteration cap drift apart unseen), so the decision is checked
    against FP64 arithmetic from the device's captured cost and warm start:
    at the check where the earlier side stopped, the FP64 marginal error of
    the device's iterates is within the FP32 floor n 2^-24 (4 + max|C| / reg
    log2 e) of the threshold on the side the device decided (at or below
    the threshold where the device stopped, at or above where it went on)."""
    counts = [p.iteration_count for p in device_profiles]
    for k, (a, b) in enumerate(zip(counts, run.sinkhorn_iterations)):
        if a == b:
            continue
        C, u0, v0 = states[k]
        floor = n * 2.0 ** -24 * (4.0 + float(C.abs().max()) / reg * 1.4426950408889634)
        stop = min(a, b)  # the earlier stop's iteration count: its check was at iteration stop - 1
        index = (stop - 1) // eval_freq - 1  # the LOG checks are at iterations 10, 20, ...
        st = oracle.SinkhornState(C.shape[0], "LOG")
        st.u, st.v = u0.double().clone(), v0.double().clone()
        _, _, _, res = st.solve_log(C.double(), reg, stop, 0.0, eval_freq)
        err64 = res.errors[index]

        def at(errors):
            return f"{errors[index]:.4e}" if len(errors) > index else "not recorded"

        print(f"first differing solve {k}: device {a}, oracle {b} iterations; at iteration {stop - 1} the device's "
              f"error {at(device_profiles[k].errors)}, FP64 from the device's state {err64:.4e}, oracle's own "
              f"{at(run.sinkhorn_errors[k])}; threshold {threshold:g}, FP32 floor {floor:.2e}")
        if a < b:
            assert err64 <= threshold + floor, (k, err64, floor)
        else:
            assert err64 >= threshold - floor, (k, err64, floor)
        return k
    return None


def _capture_solves(monkeypatch, profiles=None):
    """Record each LOG solve's cost and incoming potentials (on the host),
    and with `profiles` each solve's SinkhornProfile."""
    from cugal_amd import sinkhorn as sk

    states = []
    solve_log = sk.SinkhornState.solve_log

    def spy(self, C, config, profile=None):
        states.append((C.detach().double().cpu(), self.u.detach().double().cpu(), self.v.detach().double().cpu()))
        out = solve_log(self, C, config, profile)
        if profiles is not None:
            profiles.append(profile)
        return out

    monkeypatch.setattr(sk.SinkhornState, "solve_log", spy)
    return states


def test_reference_dataset_published_configuration_within_half_a_point(device, monkeypatch):
    # in-arenas (1133 nodes, 5400 edges, one self-loop; the reference's
    # benchmark/data file, copied to tests/golden) with 10 % of source edges
    # removed (seed 283841238) and the published configuration, end to end:
    # node accuracy within 0.5 pp of the FP64 oracle, the same Sinkhorn
    # iteration count for every one of the 150 solves.  (The FP32 CPU
    # restatement agrees with FP64 here to 0.18 pp with equal counts: this
    # instance is well conditioned.)
    import copy
    import os

    from cugal_amd import cugal
    from cugal_amd.profile import Profile

    golden = os.path.join(os.path.dirname(__file__), "golden")
    gen = np.random.default_rng(seed=283841238)
    s, t = Graph(GraphKind.IN_ARENAS, {}).get(gen, golden)
    source, target, mapping, _ = generate_graph(s, t, gen, NoiseLevel(0.1, 0.0, False))
    profile = Profile()
    states = _capture_solves(monkeypatch)
    _, answer = cugal(copy.deepcopy(source), copy.deepcopy(target), _published_config(True), profile)
    run = _oracle_run(source, target, _published_config(False))
    acc = float(np.mean(np.array([m for _, m in answer]) == np.asarray(mapping)))
    ref_acc = float(np.mean(np.array([m for _, m in run.mapping]) == np.asarray(mapping)))
    counts = [p.iteration_count for p in profile.sinkhorn_profiles]
    print(f"\nin-arenas, published configuration: accuracy {acc:.4f} vs FP64 oracle {ref_acc:.4f}; Sinkhorn "
          f"iterations {sum(counts)} vs {sum(run.sinkhorn_iterations)}")
    first = _first_count_difference(profile.sinkhorn_profiles, run, states, 1133, 0.3, 1e-3)
    assert first is None
    assert abs(acc - ref_acc) <= 0.005 + 1e-12


def test_harness_on_a_sensitive_reference_dataset(device, monkeypatch):
    # ca-netscience (379 nodes, 914 edges; the reference's benchmark/data
    # file, copied to tests/golden) with the published configuration and the
    # first noise level of the reference's run (10 % of source edges removed;
    # seed 283841238; four instances, as the published experiment averages
    # four runs), through the experiment harness.  A sensitivity check, not
    # the parity gate (in-arenas above holds 0.5 pp with equal counts in every
    # solve): on this graph a large share of the rows of Q have their two
    # largest entries within 1e-4 of each other (measured 43-45 % in the FP64
    # oracle's Q), so rounding decides which way those rows fall, and once
    # two runs part, their Frank-Wolfe steps settle on different
    # near-permutations.  What the test holds, per instance: where the device
    # and the oracle part (the first solve whose Sinkhorn count differs), the
    # device's stop is a threshold crossing within FP32 rounding of its own
    # iterates, against FP64 from its captured state; and the tied share of
    # rows is large.  The FP64 oracle run with the device's own per-solve
    # counts is also reported: its accuracy barely moves from the oracle's own
    # (measured on 3 of 4 instances unchanged), so the accuracy gap is in
    # the tied rows, not in the stop decisions.  Node accuracy and structure
    # scores are held to 3 pp on the means over the four instances, the size
    # of the measured spread (single instances part by up to 4 pp either
    # way; the FP64 oracle alone gives 0.6201 on one host and 0.6016 on
    # another for the first instance).  (The published 0.642 comes from a
    # build with the LOG marginal-check quirk; context, not a bound.)
    import os

    golden = os.path.join(os.path.dirname(__file__), "golden")
    seed = 283841238
    level = NoiseLevel(0.1, 0.0, False)
    runs = 4
    exp = Experiment(algorithms=[Algorithm(_published_config(True), False)],
                     graphs=[Graph(GraphKind.CA_NETSCIENCE, {})], noise_levels=[level], seed=seed, num_runs=runs,
                     data_dir=golden)
    profiles = []
    states = _capture_solves(monkeypatch, profiles)
    result = exp.run().results[0][0][0]  # the harness's mean over the runs
    assert len(states) == len(profiles) == 150 * runs

    gen = np.random.default_rng(seed=seed)
    s, t = Graph(GraphKind.CA_NETSCIENCE, {}).get(gen, golden)
    assert (s.number_of_nodes(), s.number_of_edges()) == (379, 914)
    pairs = [generate_graph(s, t, gen, level) for _ in range(runs)]
    ref_accs, forced_accs, ref_scores = [], [], []
    for r, (source, target, mapping, _) in enumerate(pairs):
        k = slice(150 * r, 150 * (r + 1))
        run = _oracle_run(source, target, _published_config(False))
        answer = np.array([m for _, m in run.mapping])
        ref_accs.append(float(np.mean(answer == np.asarray(mapping))))
        ref_scores.append(_oracle_scores(source, target, answer, mapping))
        _first_count_difference(profiles[k], run, states[k], 379, 0.3, 1e-3)
        counts = [p.iteration_count for p in profiles[k]]
        forced = _oracle_run(source, target, _published_config(False, forced_iterations=counts))
        assert forced.sinkhorn_iterations == counts
        forced_accs.append(float(np.mean(np.array([m for _, m in forced.mapping]) == np.asarray(mapping))))
        top2 = forced.P.topk(2, dim=1).values
        tied = float(((top2[:, 0] - top2[:, 1]) <= 1e-4 * top2[:, 0]).double().mean())
        print(f"\ninstance {r}: FP64 oracle {ref_accs[-1]:.4f}, with the device's counts {forced_accs[-1]:.4f}; "
              f"rows of its Q tied within 1e-4: {tied:.1%}")
        assert tied >= 0.25
    ref_acc = float(np.mean(ref_accs))
    ref_mean = np.mean(np.array(ref_scores, dtype=np.float64), axis=0)
    print(f"ca-netscience, 10 % source noise, mean of {runs} instances: accuracy {result.accuracy:.4f} (device "
          f"harness) vs {ref_acc:.4f} (FP64 oracle; with the device's counts {np.mean(forced_accs):.4f}); "
          f"ICS/EC/S3 {result.ics:.4f}/{result.ec:.4f}/{result.sss:.4f} vs "
          f"{ref_mean[0]:.4f}/{ref_mean[1]:.4f}/{ref_mean[2]:.4f}")
    assert abs(result.accuracy - ref_acc) <= 0.03
    for got, want in zip((result.ics, result.ec, result.sss), ref_mean):
        assert abs(got - want) <= 0.03
    assert result.profile.frank_wolfe_iterations == 150


@pytest.mark.parametrize("sparse", [True, False])
def test_dataset_with_a_self_loop_quasi_permutation_matches_oracle(device, sparse):
    # in-arenas (1133 nodes, 5400 edges, one self-loop at 565; the
    # reference's benchmark/data file, copied to tests/golden) with 10 % of
    # source edges removed: the loop stays in both graphs.  With a zero
    # threshold every solve runs its full count, so Q follows the FP64
    # oracle entry by entry: the sparse path counts the loop twice, the
    # dense path once, and the features are the GPU path's.
    import copy
    import os

    from cugal_amd import cugal
    from cugal_amd.profile import Profile

    golden = os.path.join(os.path.dirname(__file__), "golden")
    gen = np.random.default_rng(seed=283841238)
    s, t = Graph(GraphKind.IN_ARENAS, {}).get(gen, golden)
    assert (s.number_of_nodes(), s.number_of_edges(), nx.number_of_selfloops(s)) == (1133, 5400, 1)
    source, target, mapping, _ = generate_graph(s, t, gen, NoiseLevel(0.1, 0.0, False))
    assert nx.number_of_selfloops(source) == nx.number_of_selfloops(target) == 1
    cfg = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=0.3, sinkhorn_iterations=30, sinkhorn_threshold=0.0, mu=2.0,
                 iter_count=2, frank_wolfe_iter_count=3, use_sparse_adjacency=sparse,
                 hungarian_method=HungarianMethod.SCIPY)
    ref = oracle.OracleConfig(sinkhorn_regularization=0.3, sinkhorn_method="LOG", sinkhorn_threshold=0.0, mu=2.0,
                              iter_count=2, frank_wolfe_iter_count=3, use_sparse_adjacency=sparse,
                              hungarian_method="SCIPY")
    ref.sinkhorn_iterations = 30
    P, answer = cugal(copy.deepcopy(source), copy.deepcopy(target), cfg, Profile())
    run = oracle.align(copy.deepcopy(source), copy.deepcopy(target), ref)
    R = run.P.numpy()
    # The feature distances reach 164 here (mu 2), so |C|/reg reaches ~546
    # and each FP32 exponent carries up to 2^-24 (4 + 546 log2 e) = 4.7e-5
    # relative rounding (measured: 5.4e-5 on 7 of 1.28M entries); Q is held
    # to 1e-4 relative, against 1e-5 for the synthetic cases (|C/reg| <= 33).
    np.testing.assert_allclose(P, R, rtol=1e-4, atol=1e-6 * float(R.max()))


def test_sensitive_dataset_quasi_permutation_matches_oracle(device):
    # ca-netscience (the graph of the sensitivity check above) with the
    # published sparse configuration and a zero threshold: every solve runs
    # its full count, so no stop decision can part device and oracle, and Q
    # follows the FP64 oracle entry by entry -- the near-ties that decide
    # the node accuracy there lie in the LAP's reading of Q, not in Q.  Held
    # to the in-arenas bound: |C|/reg reaches a few hundred, so each FP32
    # exponent carries up to ~5e-5 relative rounding.
    import copy
    import os

    from cugal_amd import cugal
    from cugal_amd.profile import Profile

    golden = os.path.join(os.path.dirname(__file__), "golden")
    gen = np.random.default_rng(seed=283841238)
    s, t = Graph(GraphKind.CA_NETSCIENCE, {}).get(gen, golden)
    assert (s.number_of_nodes(), s.number_of_edges()) == (379, 914)
    source, target, mapping, _ = generate_graph(s, t, gen, NoiseLevel(0.1, 0.0, False))
    cfg = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=0.3, sinkhorn_iterations=30, sinkhorn_threshold=0.0, mu=2.0,
                 iter_count=2, frank_wolfe_iter_count=3, use_sparse_adjacency=True,
                 hungarian_method=HungarianMethod.SCIPY)
    ref = oracle.OracleConfig(sinkhorn_regularization=0.3, sinkhorn_method="LOG", sinkhorn_threshold=0.0, mu=2.0,
                              iter_count=2, frank_wolfe_iter_count=3, use_sparse_adjacency=True,
                              hungarian_method="SCIPY")
    ref.sinkhorn_iterations = 30
    P, answer = cugal(copy.deepcopy(source), copy.deepcopy(target), cfg, Profile())
    run = oracle.align(copy.deepcopy(source), copy.deepcopy(target), ref)
    R = run.P.numpy()
    rel = float(np.abs(P - R).max() / R.max())
    print(f"\nca-netscience, fixed iterations: max |Q - Q_FP64| / max Q = {rel:.2e}")
    np.testing.assert_allclose(P, R, rtol=1e-4, atol=1e-6 * float(R.max()))


def test_published_pairs_match_the_oracle_fixture(device, monkeypatch):
    # The reference's published experiment pairs (the shipped graphs, seed
    # 283841238, drawn in the order of benchmark/experiment.py:440-460) at
    # source noise 0.1 with the published Config, against the FP64 oracle's
    # alignment of the same pairs (tests/golden/published_pairs_oracle.json,
    # written by tests/golden/published_pairs_oracle.py).  The published
    # accuracies themselves are not a bar: they come from a source version
    # other than the one the oracle restates (with the in-place marginal
    # check of cugal/sinkhorn.py:67-74 no solve converges; see the fixture's
    # "quirk" entry).  Per graph, over the four runs: the Sinkhorn iteration
    # total within 1 % and the mean node accuracy within 0.5 pp; in every run
    # whose counts differ, the first differing solve is a threshold crossing
    # within FP32 rounding of the device's own state.
    import copy
    import importlib.util
    import os
    from types import SimpleNamespace

    from cugal_amd import cugal
    from cugal_amd.profile import Profile
    from cugal_amd.synthetic import alignment_accuracy, answer_columns

    golden = os.path.join(os.path.dirname(__file__), "golden")
    spec = importlib.util.spec_from_file_location("published_pairs_oracle",
                                                  os.path.join(golden, "published_pairs_oracle.py"))
    gen_mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen_mod)
    fixture = json.load(open(os.path.join(golden, "published_pairs_oracle.json")))["graphs"]
    pairs = gen_mod.published_pairs()
    states = _capture_solves(monkeypatch)
    checked = 0
    for name, runs in sorted(fixture.items()):
        if name.endswith(" quirk"):
            continue
        accs, refs, its, ref_its = [], [], 0, 0
        for r, want in sorted(runs.items()):
            source, target, mapping, _ = pairs[name][int(r)]
            profile = Profile()
            states.clear()
            _, answer = cugal(copy.deepcopy(source), copy.deepcopy(target), _published_config(True), profile)
            run = SimpleNamespace(sinkhorn_iterations=want["sinkhorn_iterations"],
                                  sinkhorn_errors=[[] for _ in want["sinkhorn_iterations"]])
            n = max(source.number_of_nodes(), target.number_of_nodes())
            _first_count_difference(profile.sinkhorn_profiles, run, states, n, 0.3, 1e-3)
            accs.append(alignment_accuracy(mapping, answer_columns(answer)))
            refs.append(want["accuracy"])
            its += sum(p.iteration_count for p in profile.sinkhorn_profiles)
            ref_its += sum(want["sinkhorn_iterations"])
        acc, ref = float(np.mean(accs)), float(np.mean(refs))
        print(f"\n{name} published pairs ({len(accs)} runs): accuracy {acc:.4f} vs FP64 oracle {ref:.4f} "
              f"(per run {np.round(accs, 4).tolist()} vs {np.round(refs, 4).tolist()}); Sinkhorn iterations "
              f"{its} vs {ref_its}")
        assert abs(its - ref_its) <= 0.01 * ref_its
        assert abs(acc - ref) <= 0.005 + 1e-12
        checked += 1
    assert checked >= 1


def test_published_pair_at_full_width_quasi_permutation_matches_oracle(device):
    # bio-dmela, the largest published graph (n = 7393: an odd width, so the
    # fused sweeps take their any-width path), published pair 0 at the
    # published regularisation and mu, two Frank-Wolfe steps of 20 Sinkhorn
    # iterations each (fixed counts: threshold 0): Q entry by entry against
    # the FP64 oracle, held to the in-arenas bound (|C|/reg reaches a few
    # hundred, so each FP32 exponent carries up to ~5e-5 relative rounding).
    import copy
    import importlib.util
    import os

    from cugal_amd import cugal
    from cugal_amd.profile import Profile

    golden = os.path.join(os.path.dirname(__file__), "golden")
    spec = importlib.util.spec_from_file_location("published_pairs_oracle",
                                                  os.path.join(golden, "published_pairs_oracle.py"))
    gen_mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen_mod)
    source, target, _, _ = gen_mod.published_pairs()["BIO_DMELA"][0]
    assert max(source.number_of_nodes(), target.number_of_nodes()) == 7393
    cfg = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=0.3, sinkhorn_iterations=20, sinkhorn_threshold=0.0, mu=2.0,
                 iter_count=1, frank_wolfe_iter_count=2, use_sparse_adjacency=True,
                 hungarian_method=HungarianMethod.SCIPY)
    ref = oracle.OracleConfig(sinkhorn_regularization=0.3, sinkhorn_method="LOG", sinkhorn_threshold=0.0, mu=2.0,
                              iter_count=1, frank_wolfe_iter_count=2, use_sparse_adjacency=True,
                              hungarian_method="SCIPY")
    ref.sinkhorn_iterations = 20
    profile = Profile()
    P, _ = cugal(copy.deepcopy(source), copy.deepcopy(target), cfg, profile)
    assert [p.iteration_count for p in profile.sinkhorn_profiles] == [20, 20]
    run = _oracle_run(source, target, ref, threads=16)
    R = run.P.numpy()
    rel = float(np.abs(P - R).max() / R.max())
    print(f"\nbio-dmela (n = 7393), fixed iterations: max |Q - Q_FP64| / max Q = {rel:.2e}")
    np.testing.assert_allclose(P, R, rtol=1e-4, atol=1e-6 * float(R.max()))
