"""Seeded synthetic alignment instances: the noise and permutation model of
benchmark/experiment.py:35-112 and the accuracy of :143-144.

A source graph is generated with networkx, its edges are relabelled by a
seeded permutation to give the target, a fraction of target edges is removed
(never leaving an endpoint isolated) and, with refill, the same number of
random non-edges is added.  ``source_mapping[i]`` is the target node of source
node i, so the accuracy of an alignment is mean(answer == source_mapping).
"""

from __future__ import annotations

from dataclasses import dataclass

import networkx as nx
import numpy as np


@dataclass(frozen=True)
class NoiseLevel:
    """benchmark/experiment.py:23-32."""

    source_noise: float
    target_noise: float
    refill_edges: bool


def refill_edges(edges: np.ndarray, dimension: int, edge_amount: int, generator: np.random.Generator) -> np.ndarray:
    """Add `edge_amount` random non-edges (benchmark/experiment.py:35-47)."""
    if edge_amount == 0:
        return edges
    edge_set = {tuple(row) for row in np.sort(edges).tolist()}
    new_edges = []
    while len(new_edges) < edge_amount:
        new_edge = generator.integers(0, dimension, size=2)
        sorted_new_edge = tuple(np.sort(new_edge).tolist())
        if sorted_new_edge not in edge_set and new_edge[0] != new_edge[1]:
            edge_set.add(sorted_new_edge)
            new_edges.append(new_edge)
    return np.append(edges, new_edges, axis=0)


def remove_edges(edges: np.ndarray, noise: float, generator: np.random.Generator) -> np.ndarray:
    """Drop each edge with probability `noise` unless an endpoint would lose
    its last edge (benchmark/experiment.py:50-60)."""
    bin_count = np.bincount(edges.flatten())
    # One draw per edge, as the reference's per-edge generator.random().
    draws = generator.random(len(edges))
    rows_to_delete = []
    for i in np.nonzero(draws < noise)[0]:
        e, f = edges[i]
        if bin_count[e] > 1 and bin_count[f] > 1:
            bin_count[e] -= 1
            bin_count[f] -= 1
            rows_to_delete.append(i)
    return np.delete(edges, rows_to_delete, axis=0)


def permute_edges(edges: np.ndarray, dimension: int, seed: int = 0):
    """Relabel by a seeded permutation (benchmark/experiment.py:71-78 fixes
    the seed at 0)."""
    source_permutation = np.arange(dimension)
    target_permutation = np.random.RandomState(seed=seed).permutation(dimension)
    source_mapping = source_permutation[target_permutation.argsort()]
    target_mapping = target_permutation[source_permutation.argsort()]
    return source_mapping[edges], source_mapping, target_mapping


@dataclass
class Instance:
    source: nx.Graph
    target: nx.Graph
    source_mapping: np.ndarray
    description: dict


@dataclass
class EdgeInstance:
    """An instance kept as edge arrays: (E, 2) int64 pairs over nodes
    0..n-1.  Large instances skip networkx (building a 2-million-edge graph
    object costs tens of seconds); ``adjacencies`` builds the CSR of both
    graphs on the device straight from the pairs, the same CSR as
    ``Adjacency.from_graph`` of the padded graphs."""

    source_edges: np.ndarray
    target_edges: np.ndarray
    n: int
    source_mapping: np.ndarray
    description: dict

    def graphs(self) -> Instance:
        source = nx.from_edgelist(self.source_edges.tolist())
        target = nx.from_edgelist(self.target_edges.tolist())
        return Instance(source, target, self.source_mapping, dict(self.description))

    def adjacencies(self, device):
        import torch

        from .adjacency import Adjacency, edges_to_csr

        out = []
        for edges in (self.source_edges, self.target_edges):
            t = torch.from_numpy(np.ascontiguousarray(edges, dtype=np.int32)).to(device)
            col, rp = edges_to_csr(t, self.n, symmetrize=True)
            out.append(Adjacency(col, rp))
        return tuple(out)


def make_edge_instance(source_edges: np.ndarray, noise: NoiseLevel, seed: int = 0, permutation_seed: int = 0,
                       description: dict | None = None) -> EdgeInstance:
    """generate_graph + add_synthetic_noise (benchmark/experiment.py:81-112,
    :315-324) on edge arrays.  The generator draws in the reference's order:
    source removal, target removal, source refill, target refill.  The draws
    of remove_edges are taken as one vector per call, the same stream
    values."""
    generator = np.random.default_rng(seed)
    source_edges = np.asarray(source_edges, dtype=np.int64)
    if np.amin(source_edges) != 0:
        source_edges = source_edges - np.amin(source_edges)
    dimension = int(np.amax(source_edges)) + 1
    target_edges, source_mapping, _ = permute_edges(source_edges, dimension, permutation_seed)
    edge_count = source_edges.shape[0]
    source_edges = remove_edges(source_edges, noise.source_noise, generator)
    target_edges = remove_edges(target_edges, noise.target_noise, generator)
    if noise.refill_edges:
        source_edges = refill_edges(source_edges, dimension, edge_count - source_edges.shape[0], generator)
        target_edges = refill_edges(target_edges, dimension, edge_count - target_edges.shape[0], generator)
    return EdgeInstance(source_edges, target_edges, dimension, source_mapping, dict(description or {}))


def make_instance(graph: nx.Graph, noise: NoiseLevel, seed: int = 0, permutation_seed: int = 0,
                  description: dict | None = None) -> Instance:
    """make_edge_instance on the edges of a networkx graph, returned as
    networkx graphs."""
    return make_edge_instance(np.array(graph.edges), noise, seed, permutation_seed, description).graphs()


def alignment_accuracy(ground_truth: np.ndarray, mapping) -> float:
    """benchmark/experiment.py:143-144."""
    return float(np.mean(np.asarray(mapping) == np.asarray(ground_truth)))


def answer_columns(mapping: list[tuple[int, int]]) -> np.ndarray:
    return np.array([x for _, x in mapping])


# ---- the workloads ---------------------------------------------------------

def erdos_renyi(n: int, p: float, noise: float, seed: int = 0) -> Instance:
    g = nx.erdos_renyi_graph(n, p, seed=seed)
    return make_instance(g, NoiseLevel(0.0, noise, True), seed,
                         description={"graph": "erdos_renyi", "n": n, "p": p, "noise": noise, "seed": seed})


def barabasi_albert(n: int, m: int, noise: float, seed: int = 0) -> Instance:
    g = nx.barabasi_albert_graph(n, m, seed=seed)
    return make_instance(g, NoiseLevel(0.0, noise, True), seed,
                         description={"graph": "barabasi_albert", "n": n, "m": m, "noise": noise, "seed": seed})


def watts_strogatz(n: int, k: int, p: float, noise: float, seed: int = 0) -> Instance:
    g = nx.watts_strogatz_graph(n, k, p, seed=seed)
    return make_instance(g, NoiseLevel(0.0, noise, True), seed,
                         description={"graph": "watts_strogatz", "n": n, "k": k, "p": p, "noise": noise,
                                      "seed": seed})


def stochastic_block_edges(blocks: int, block_size: int, deg_in: float, deg_out: float,
                           seed: int = 0) -> np.ndarray:
    """Edges of a stochastic block model with `blocks` equal blocks: every
    pair inside a block is an edge with probability p_in = deg_in /
    (block_size - 1), every pair across two blocks with p_out = deg_out /
    (n - block_size), independently.  Per block pair the edge count is drawn
    binomially and that many distinct pairs uniformly (the same
    distribution), all from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    n = blocks * block_size
    p_in = deg_in / (block_size - 1)
    p_out = deg_out / (n - block_size)
    iu, ju = np.triu_indices(block_size, 1)
    parts = []
    for a in range(blocks):
        base = a * block_size
        m = rng.binomial(len(iu), p_in)
        pick = rng.choice(len(iu), size=m, replace=False)
        parts.append(np.stack((base + iu[pick], base + ju[pick]), axis=1))
        for b in range(a + 1, blocks):
            m = rng.binomial(block_size * block_size, p_out)
            pick = rng.choice(block_size * block_size, size=m, replace=False)
            parts.append(np.stack((base + pick // block_size, b * block_size + pick % block_size), axis=1))
    return np.concatenate(parts).astype(np.int64)


def stochastic_block(blocks: int, block_size: int, deg_in: float, deg_out: float, noise: float,
                     seed: int = 0) -> EdgeInstance:
    """Stochastic block model instance (stochastic_block_edges), 1 % edge
    noise with refill and a seeded permutation, as edge arrays."""
    edges = stochastic_block_edges(blocks, block_size, deg_in, deg_out, seed)
    return make_edge_instance(edges, NoiseLevel(0.0, noise, True), seed,
                              description={"graph": "stochastic_block", "blocks": blocks, "block_size": block_size,
                                           "deg_in": deg_in, "deg_out": deg_out, "noise": noise, "seed": seed})
