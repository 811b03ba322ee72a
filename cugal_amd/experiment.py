"""Experiment harness: the structure and on-disk formats of the reference's
benchmark/experiment.py.

* ``create_graph_from_str`` reads SNAP edge lists (experiment.py:155-163);
  ``GraphKind`` / ``Graph`` name the reference's datasets and synthetic
  families (:166-294), loading the files from a caller-given data folder;
* ``generate_graph`` adds the seeded permutation and edge noise (:297-324);
* ``Result`` holds ICS, EC, S3, accuracy and the profile (:327-392); the
  three structure scores come from the device kernel ``cg_alignment_counts``
  over CSRs in the graphs' node order (the rows nx.to_numpy_array gives the
  reference's dense matrices), so they are the reference's numbers without an
  n x n matrix;
* ``Experiment`` runs graphs x noise levels x algorithms (:404-501) through
  ``cugal_amd.cugal``; ``ExperimentResults`` writes and reads the results
  JSON of :515-560 (same keys, so result files are interchangeable).

The original FUGAL implementation (``Algorithm.use_fugal``) and the HTTP
download of ca-HepTh are not provided and raise NotImplementedError.
"""

from __future__ import annotations

import dataclasses
import datetime
import json
import os
import subprocess
from dataclasses import dataclass
from enum import Enum
from typing import Iterable

import networkx as nx
import numpy as np
import torch

from . import _native as nat
from .adjacency import Adjacency, edges_to_csr
from .config import Config
from .pred import cugal
from .profile import Profile
from .synthetic import NoiseLevel as _NoiseLevel
from .synthetic import permute_edges, refill_edges, remove_edges


@dataclass(frozen=True)
class NoiseLevel(_NoiseLevel):
    """benchmark/experiment.py:22-32."""

    def __str__(self):
        return str(dataclasses.asdict(self))

    def to_dict(self) -> dict:
        return dataclasses.asdict(self)


def create_graph_from_str(file: str) -> nx.Graph:
    """A graph from SNAP edge-list text: one `u v` pair per line, lines
    starting with '#' ignored, lines without exactly two integers skipped
    (benchmark/experiment.py:164-172)."""
    edges = [tuple(map(int, line.split())) for line in file.split("\n") if not line.startswith("#")]
    return nx.from_edgelist([edge for edge in edges if len(edge) == 2])


def load_edge_list(path: str) -> nx.Graph:
    with open(path, "r") as f:
        return create_graph_from_str(f.read())


class GraphKind(Enum):
    """benchmark/experiment.py:166-204."""

    CA_HEP = "CA_HEP"
    CA_ERDOS = "CA_ERDOS"
    CA_GRQC = "CA_GRQC"
    CA_NETSCIENCE = "CA_NETSCIENCE"
    CA_ASTROPH = "CA_ASTROPH"
    CIT_HEP_PH = "CIT_HEP_PH"
    BIO_DMELA = "BIO_DMELA"
    BIO_CELEGANS = "BIO_CELEGANS"
    IN_ARENAS = "IN_ARENAS"
    INF_POWER = "INF_POWER"
    INF_EUROROAD = "INF_EUROROAD"
    EMAIL_ENRON = "EMAIL_ENRON"
    SOC_FACEBOOK = "SOC_FACEBOOK"
    SOC_HAMSTERSTER = "SOC_HAMSTERSTER"
    SOCFB_BOWDOIN47 = "SOCFB_BOWDOIN47"
    SOCFB_HAMILTON46 = "SOCFB_HAMILTON46"
    SOCFB_HAVERFORD76 = "SOCFB_HAVERFORD76"
    SOCFB_SWARTHMORE42 = "SOCFB_SWARTHMORE42"
    NEWMAN_WATTS = "NEWMAN_WATTS"
    LOBSTER = "LOBSTER"
    PREDEFINED_GRAPHS = "PREDEFINED_GRAPHS"
    AVERAGED = "AVERAGED"


# File of each real-world kind in the data folder, as the reference opens
# them (benchmark/experiment.py:232-289; SOC_FACEBOOK and SOC_HAMSTERSTER
# read the socfb files there, and so they do here).
DATA_FILES = {
    GraphKind.IN_ARENAS: "in-arenas.txt",
    GraphKind.INF_POWER: "inf-power.txt",
    GraphKind.BIO_DMELA: "bio-dmela.txt",
    GraphKind.BIO_CELEGANS: "bio-celegans.txt",
    GraphKind.CA_NETSCIENCE: "ca-netscience.txt",
    GraphKind.CA_ERDOS: "ca-Erdos992.txt",
    GraphKind.CA_GRQC: "ca-GrQc.txt",
    GraphKind.CA_ASTROPH: "CA-AstroPh.txt",
    GraphKind.CIT_HEP_PH: "cit-HepPh.txt",
    GraphKind.INF_EUROROAD: "inf-euroroad.txt",
    GraphKind.EMAIL_ENRON: "email-Enron.txt",
    GraphKind.SOC_FACEBOOK: "socfb-Bowdoin47.txt",
    GraphKind.SOC_HAMSTERSTER: "socfb-Hamilton46.txt",
    GraphKind.SOCFB_BOWDOIN47: "socfb-Bowdoin47.txt",
    GraphKind.SOCFB_HAMILTON46: "socfb-Hamilton46.txt",
    GraphKind.SOCFB_HAVERFORD76: "socfb-Haverford76.txt",
    GraphKind.SOCFB_SWARTHMORE42: "socfb-Swarthmore42.txt",
}


@dataclass(frozen=True)
class Graph:
    """benchmark/experiment.py:207-294.  Real-world kinds are read from
    `data_dir` (the reference reads the relative folder "data")."""

    kind: GraphKind
    parameters: dict

    def get(self, generator: np.random.Generator, data_dir: str = "data") -> tuple[nx.Graph, nx.Graph | None]:
        if self.kind == GraphKind.NEWMAN_WATTS:
            return nx.newman_watts_strogatz_graph(**self.parameters, seed=generator), None
        if self.kind == GraphKind.LOBSTER:
            return nx.random_lobster(**self.parameters, seed=generator), None
        if self.kind == GraphKind.PREDEFINED_GRAPHS:
            return load_edge_list(self.parameters["source_file"]), load_edge_list(self.parameters["target_file"])
        if self.kind in DATA_FILES:
            return load_edge_list(os.path.join(data_dir, DATA_FILES[self.kind])), None
        raise NotImplementedError(f"graph kind {self.kind.value} is not available offline")

    def to_dict(self) -> dict:
        return {"kind": self.kind.value, "parameters": self.parameters}

    def __str__(self):
        return self.kind.value + " " + str(self.parameters)

    @classmethod
    def from_dict(cls, d: dict):
        d = dict(d)
        d["kind"] = GraphKind[d["kind"]]
        return cls(**d)


def _zero_based(edges: np.ndarray) -> np.ndarray:
    return edges - np.amin(edges) if np.amin(edges) != 0 else edges


def generate_graph(source: nx.Graph, target: nx.Graph | None, generator: np.random.Generator,
                   noise_level: NoiseLevel) -> tuple[nx.Graph, nx.Graph, np.ndarray, np.ndarray]:
    """benchmark/experiment.py:297-324 with add_synthetic_noise (:81-112):
    returns (source, target, source_mapping, target_mapping)."""
    if target is None:
        source_edges = _zero_based(np.array(source.edges))
        dimension = int(np.amax(source_edges)) + 1
        target_edges, source_mapping, target_mapping = permute_edges(source_edges, dimension)
        edge_count = source_edges.shape[0]
        source_edges = remove_edges(source_edges, noise_level.source_noise, generator)
        target_edges = remove_edges(target_edges, noise_level.target_noise, generator)
        if noise_level.refill_edges:
            source_edges = refill_edges(source_edges, dimension, edge_count - source_edges.shape[0], generator)
            target_edges = refill_edges(target_edges, dimension, edge_count - target_edges.shape[0], generator)
        return (nx.from_edgelist(source_edges.tolist()), nx.from_edgelist(target_edges.tolist()), source_mapping,
                target_mapping)
    source_edges = _zero_based(np.array(source.edges))
    target_edges = _zero_based(np.array(target.edges))
    target_edges, source_mapping, target_mapping = permute_edges(target_edges, int(np.amax(target_edges)) + 1)
    return (nx.from_edgelist(source_edges.tolist()), nx.from_edgelist(target_edges.tolist()), source_mapping,
            target_mapping)


# ---- structure scores ------------------------------------------------------

def _score_adjacency(g, device: torch.device, by_label: bool) -> Adjacency:
    """CSR of a networkx graph whose row r is its r-th node (the order
    nx.to_numpy_array gives the reference's matrices) or, with `by_label`,
    node r (labels must be 0..n-1); or of a dense 0/1 matrix."""
    if isinstance(g, nx.Graph):
        if by_label:
            if set(g.nodes()) != set(range(g.number_of_nodes())):
                raise ValueError("standard scores need graphs labelled 0..n-1")
            pos = {v: v for v in g.nodes()}
        else:
            pos = {v: i for i, v in enumerate(g.nodes())}
        pairs = np.array([(pos[u], pos[v]) for u, v in g.edges()], dtype=np.int32).reshape(-1, 2)
        if g.is_directed():
            pairs = pairs[:, ::-1].copy()  # edge u -> v is the entry [u][v]
        edges = torch.from_numpy(pairs).to(device)
        col, rp = edges_to_csr(edges, g.number_of_nodes(), symmetrize=not g.is_directed())
        return Adjacency(col, rp)
    dense = torch.as_tensor(np.asarray(g), device=device)
    if dense.dim() != 2 or dense.shape[0] != dense.shape[1]:
        raise ValueError("adjacency must be a square matrix or a networkx graph")
    if bool(((dense != 0) & (dense != 1)).any()):
        raise ValueError("the structure scores need a 0/1 adjacency")
    return Adjacency.from_dense(dense)


def _ratio(a: int, b: int) -> float:
    # numpy's semantics for the reference's int / int: x / 0 is inf, 0 / 0 nan.
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.divide(np.float64(a), np.float64(b)))


def structure_scores(source, target, answer, mapping, device="cuda", standard: bool = False
                     ) -> tuple[float, float, float]:
    """(ICS, EC, S3) of benchmark/experiment.py:115-141 as Result.calculate
    calls them: `answer` = computed target node per source node, `mapping` =
    ground truth.  `source`/`target` are networkx graphs or dense 0/1
    matrices.  Returns zeros where the reference's numpy code raises (answers
    out of range, lengths that differ).

    The reference compares S[answer][:, answer] with T[mapping][:, mapping],
    the matrices' rows in the graphs' node order, so the true alignment
    scores 1 only when the ground truth is the identity; that is what it
    reports and what this returns by default.  `standard=True` gives the
    usual definitions instead, on label-indexed adjacencies: EC = conserved
    edges / |E_S|, ICS = conserved / edges of T induced by the answer's image,
    S3 = conserved / (|E_S| + induced - conserved), with conserved =
    #{(u, v) in E_S : (answer[u], answer[v]) in E_T}; the answer must then be
    a permutation."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise NotImplementedError("structure scores run on a ROCm device")
    answer = np.asarray(answer).reshape(-1)
    mapping = np.asarray(mapping).reshape(-1)
    if standard:
        # The reference formula with (identity, answer) on label order.
        answer, mapping = np.arange(answer.shape[0]), answer
    elif answer.shape != mapping.shape:
        return 0.0, 0.0, 0.0
    S = _score_adjacency(source, dev, standard)
    T = _score_adjacency(target, dev, standard)
    m = answer.shape[0]
    n_s, n_t = S.number_of_nodes(), T.number_of_nodes()
    if m and (answer.min() < -(2 ** 31) or answer.max() >= 2 ** 31 or mapping.min() < 0 or mapping.max() >= n_t):
        if standard:
            raise ValueError("the answer must be a permutation of the target nodes")
        return 0.0, 0.0, 0.0
    ans = torch.from_numpy(answer.astype(np.int32)).to(dev)
    truth = torch.from_numpy(mapping.astype(np.int32)).to(dev)
    counts = torch.empty(4, dtype=torch.int64, device=dev)
    ws = nat.workspace(dev, nat.query("cg_alignment_counts_workspace", n_t, m), "metrics")
    nat.call("cg_alignment_counts", nat.ptr(S.col_indices) if len(S.col_indices) else None, nat.ptr(S.row_pointers),
             n_s, nat.ptr(T.col_indices) if len(T.col_indices) else None, nat.ptr(T.row_pointers), n_t,
             nat.ptr(ans) if m else None, nat.ptr(truth) if m else None, m, nat.ptr(counts), nat.ptr(ws), ws.numel(),
             nat.stream_of(dev))
    inter, ones_s, ones_t, flags = (int(x) for x in counts.cpu().tolist())
    if flags & 2:
        raise ValueError("the " + ("answer" if standard else "ground-truth mapping")
                         + " must be a permutation of distinct target nodes")
    if flags & 1:
        return 0.0, 0.0, 0.0
    return _ratio(inter, ones_t), _ratio(inter, ones_s), _ratio(inter, ones_s + ones_t - inter)


def alignment_accuracy(ground_truth: np.ndarray, mapping: np.ndarray) -> float:
    """benchmark/experiment.py:143-144."""
    return float(np.mean(np.asarray(mapping) == np.asarray(ground_truth)))


@dataclass
class Result:
    """benchmark/experiment.py:327-392."""

    ics: float
    ec: float
    sss: float
    accuracy: float
    profile: Profile

    @classmethod
    def calculate(cls, profile: Profile, source_adjacency, target_adjacency, answer: np.ndarray, mapping: np.ndarray,
                  device="cuda"):
        ics, ec, sss = structure_scores(source_adjacency, target_adjacency, answer, mapping, device)
        return cls(ics, ec, sss, alignment_accuracy(answer, mapping), profile)

    @staticmethod
    def average(results: list["Result"]) -> "Result":
        count = len(results)
        return Result(sum(r.ics for r in results) / count, sum(r.ec for r in results) / count,
                      sum(r.sss for r in results) / count, sum(r.accuracy for r in results) / count,
                      Profile.average([r.profile for r in results]))

    def __str__(self) -> str:
        metrics = [self.ics, self.ec, self.sss, self.accuracy, self.profile.time]
        names = ["Induced Conserved Structure (ICS)", "Edge Correctness (EC)", "Symmetric Substructure Score (SSS)",
                 "Accuracy", "Time (seconds)"]
        width = max(len(name) for name in names)
        return "\n".join(f"{n:<{width}} {m}" for n, m in zip(names, metrics))

    def to_dict(self) -> dict:
        d = dataclasses.asdict(self)
        d["profile"] = self.profile.to_dict()
        return d

    @classmethod
    def from_dict(cls, d: dict):
        d = dict(d)
        d["profile"] = Profile.from_dict(d["profile"])
        return cls(**d)


@dataclass(frozen=True)
class Algorithm:
    """benchmark/experiment.py:147-163."""

    config: Config
    use_fugal: bool

    def to_dict(self) -> dict:
        return {"use_fugal": self.use_fugal, "config": self.config.to_dict()}

    def __str__(self):
        return ("Fugal" if self.use_fugal else "Cugal") + " " + str(self.config.to_dict())

    @classmethod
    def from_dict(cls, d: dict):
        return cls(config=Config.from_dict(d["config"]), use_fugal=d["use_fugal"])


def get_last_commit() -> str:
    try:
        out = subprocess.run(["git", "rev-parse", "HEAD"], capture_output=True, text=True, check=True)
        return out.stdout.strip()
    except (subprocess.CalledProcessError, OSError):
        return ""


@dataclass
class Experiment:
    """benchmark/experiment.py:404-501; `data_dir` is where the real-world
    edge lists live."""

    algorithms: list[Algorithm]
    graphs: list[Graph]
    noise_levels: list[NoiseLevel]
    debug: bool = False
    seed: int | None = None
    save_alignment: bool = False
    num_runs: int = 1
    data_dir: str = "data"

    def to_dict(self) -> dict:
        d = dataclasses.asdict(self)
        d["algorithms"] = [a.to_dict() for a in self.algorithms]
        d["graphs"] = [g.to_dict() for g in self.graphs]
        d["noise_levels"] = [dataclasses.asdict(level) for level in self.noise_levels]
        d.pop("data_dir")
        return d

    @classmethod
    def from_dict(cls, d: dict):
        d = dict(d)
        d["algorithms"] = [Algorithm.from_dict(a) for a in d["algorithms"]]
        d["graphs"] = [Graph.from_dict(g) for g in d["graphs"]]
        d["noise_levels"] = [NoiseLevel(**level) for level in d["noise_levels"]]
        return cls(**d)

    def run(self) -> "ExperimentResults":
        if self.seed is None:
            self.seed = int(np.random.randint(0xFFFFFFFF))
        assert len(self.algorithms) > 0, "no algorithms provided"
        assert len(self.noise_levels) > 0, "no noise levels provided"
        assert len(self.graphs) > 0, "no graphs provided"
        if any(a.use_fugal for a in self.algorithms):
            raise NotImplementedError("the original FUGAL implementation is not part of this library")
        generator = np.random.default_rng(seed=self.seed)
        results = []
        for graph in self.graphs:
            print(f"Running on {graph}")
            source_graph, target_graph = graph.get(generator, self.data_dir)
            graph_results = []
            for noise_level in self.noise_levels:
                print(f"Running with noise level {noise_level}")
                sources, targets, mappings = [], [], []
                for _ in range(self.num_runs):
                    source, target, source_mapping, _ = generate_graph(source_graph, target_graph, generator,
                                                                       noise_level)
                    sources.append(source)
                    targets.append(target)
                    mappings.append(source_mapping)
                noise_results = []
                for algorithm in self.algorithms:
                    print(f"Running with algorithm {algorithm}")
                    run_results = []
                    for i in range(self.num_runs):
                        profile = Profile()
                        _, answer = cugal(sources[i], targets[i], algorithm.config, profile)
                        if self.save_alignment:
                            with open(str(datetime.datetime.now()) + ".txt", "w") as f:
                                f.write("\n".join(f"{x} {y}" for x, y in answer))
                        run_results.append(Result.calculate(profile, sources[i], targets[i],
                                                            np.array([x for _, x in answer]), mappings[i],
                                                            algorithm.config.device))
                    noise_results.append(Result.average(run_results))
                graph_results.append(noise_results)
            results.append(graph_results)
        return ExperimentResults.from_results(self, results)


@dataclass
class ExperimentResults:
    """benchmark/experiment.py:515-560: results[graph][noise level][algorithm]."""

    experiment: Experiment
    commit: str
    time: str
    results: list

    def dump(self, folder: str):
        with open(os.path.join(folder, self.name() + ".json"), "w") as f:
            json.dump(self.to_dict(), f, indent=4)

    def all_results(self) -> Iterable[tuple[Graph, NoiseLevel, Algorithm, Result]]:
        for graph, graph_results in zip(self.experiment.graphs, self.results):
            for noise_level, level_results in zip(self.experiment.noise_levels, graph_results):
                for algorithm, result in zip(self.experiment.algorithms, level_results):
                    yield graph, noise_level, algorithm, result

    @classmethod
    def from_results(cls, experiment: Experiment, results: list):
        return cls(experiment=experiment, results=results, time=str(datetime.datetime.now()),
                   commit=get_last_commit())

    @classmethod
    def from_dict(cls, d: dict):
        d = dict(d)
        d["experiment"] = Experiment.from_dict(d["experiment"])
        d["results"] = [[[Result.from_dict(r) for r in b] for b in a] for a in d["results"]]
        return cls(**d)

    def name(self) -> str:
        return "-".join(g.kind.value for g in self.experiment.graphs) + "-" + self.time.replace(" ", "-")

    def to_dict(self) -> dict:
        return {
            "experiment": self.experiment.to_dict(),
            "commit": self.commit,
            "time": self.time,
            "results": [[[r.to_dict() for r in b] for b in a] for a in self.results],
        }
