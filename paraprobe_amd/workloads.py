"""The batching and serving workloads (seeded, synthetic).

Batching: 256 independent point clouds with ion counts log-uniform in
[50k, 4M], uniform or clustered frustum tips, each with its own type pair,
Rmax ~ U(1, 5) nm and kmax in {1, 10, 50}: RDF + MKNN k = 1..kmax, bin 0.01 nm.

Serving: 8 resident clustered frustum tips of 1M-50M ions and a
Poisson-arrival stream of RDF or kNN requests mixing type pairs, Rmax 1-5 nm
and k up to 100.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .capi import FAM_MKNN, FAM_RDF, Job
from . import synth


@dataclass
class BatchItem:
    n: int
    seed: int
    clustered: bool
    job: Job


def batch_spec(count: int = 256, seed: int = 3, nmin: int = 50_000, nmax: int = 4_000_000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        a, b = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        R = float(np.float32(rng.uniform(1.0, 5.0)))
        kmax = int(rng.choice([1, 10, 50]))
        job = Job([((a,), (b,))], 0.0, 0.01, R, FAM_RDF | FAM_MKNN,
                  kvalues=list(range(1, kmax + 1)))
        out.append(BatchItem(n, 1000 + i, bool(rng.integers(0, 2)), job))
    return out


def batch_clouds(spec):
    return [synth.scaled_frustum(it.n, it.seed, it.clustered) for it in spec]


SERVE_SIZES = [1_000_000, 2_000_000, 4_000_000, 7_000_000, 12_000_000, 20_000_000, 32_000_000,
               50_000_000]


def serve_tips(sizes=SERVE_SIZES, seed: int = 11):
    return [synth.scaled_frustum(n, seed + i, True) for i, n in enumerate(sizes)]


@dataclass
class Request:
    arrival: float  # seconds from the start of the stream
    cloud: int
    job: Job


def serve_stream(count: int = 10_000, rate: float = 50.0, nclouds: int = 8, seed: int = 17):
    """Poisson arrivals at `rate` per second; each request picks a cloud, a
    type pair, Rmax ~ U(1, 5) and either RDF or kNN with kmax in 1..100."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        c = int(rng.integers(0, nclouds))
        a, b = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        R = float(np.float32(rng.uniform(1.0, 5.0)))
        if rng.random() < 0.5:
            job = Job([((a,), (b,))], 0.0, 0.01, R, FAM_RDF)
        else:
            kmax = int(rng.integers(1, 101))
            job = Job([((a,), (b,))], 0.0, 0.01, R, FAM_MKNN, kvalues=list(range(1, kmax + 1)))
        out.append(Request(t, c, job))
    return out
