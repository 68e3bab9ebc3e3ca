"""Readers for PLISCuda's input surface (conf, topology, state, reactions,
constants, compartments files), mirroring the reference's C++ parsers.

* conf (PLISCuda/include/cpp/Configuration.hpp:27-144): ``key=value`` lines in
  the fixed order topology, reactions, initialState, endTime, logFreq,
  compartments, [compartmentsFile if compartments > 1], constants x C.
  File paths are relative to the working directory of the caller (as in the
  reference); ``base`` resolves them relative to another directory.
* topology (Topology.hpp:98-191): ``subvolumes: N``, blank line, then N lines
  ``v: u1 u2 ...`` in order.  Returned as CSR (the reference pads 6 slots
  with the subvolume's own id; self entries carry no hop).
* state (State.hpp:112-217): ``subvolumes: N``, ``species: S``, blank, then
  any subset of ``v: c0 .. c(S-1)``; omitted subvolumes are 0; layout [s*N+v].
* reactions (Reactions.hpp:132-223): ``reactions: R``, ``species: S``, blank,
  then lines ``r0 .. r(S-1) -> p0 .. p(S-1)``; layout [s*R + j].
* constants (cpp_utils.hpp:29-45): line 1 >= R reaction constants, line 2
  >= S diffusion constants; extra tokens ignored.
* compartments file (cpp_utils.hpp:47-56): N lines ``v: c``.

Errors raise ValueError with the reference's message text.
"""
from __future__ import annotations

import os

import numpy as np


def _split(line):
    # std::getline + split(' ') keeps empty tokens between double spaces; the
    # reference then stoi()s them and fails.  Real files use single spaces and
    # a trailing space; drop empty tokens.
    return [t for t in line.strip().split(" ") if t != ""]


def _header(line, what):
    if ":" not in line:
        raise ValueError(f"Invalid {what} number : {line.strip()}")
    try:
        return int(line.split(":", 1)[1])
    except ValueError:
        raise ValueError(f"Invalid {what} number : {line.split(':', 1)[1].strip()}") from None


def read_topology(path):
    with open(path) as f:
        lines = f.read().split("\n")
    n = _header(lines[0], "subvolumes")
    if n < 1:
        raise ValueError("Invalid: n < 1")
    rows = []
    expect = 0
    for line in lines[2:]:
        if not line.strip():
            continue
        head, _, rest = line.partition(":")
        try:
            v = int(head)
        except ValueError:
            raise ValueError(f"Invalid subvolume number : {head}") from None
        if v != expect:
            raise ValueError(f"Unexpected subvolume number. Want: {expect}Got: {head}")
        expect += 1
        nb = []
        for tok in _split(rest):
            try:
                u = int(tok)
            except ValueError:
                raise ValueError("parse of neighbour number failed") from None
            if u < 0 or u >= n:
                raise ValueError("Invalid: sub >= n")
            nb.append(u)
        rows.append(nb)
    if len(rows) != n:
        raise ValueError(f"topology lists {len(rows)} subvolumes, header says {n}")
    rp = np.zeros(n + 1, dtype=np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([u for r in rows for u in r], dtype=np.uint32)
    return {"N": n, "row_ptr": rp, "col": col}


def read_state(path):
    with open(path) as f:
        lines = f.read().split("\n")
    n = _header(lines[0], "subvolume")
    if n < 1:
        raise ValueError("Invalid: n < 1")
    s = _header(lines[1], "species")
    if s < 1:
        raise ValueError("Invalid: s < 1")
    x = np.zeros((s, n), dtype=np.int64)
    for line in lines[3:]:
        if not line.strip():
            continue
        head, _, rest = line.partition(":")
        try:
            v = int(head)
        except ValueError:
            raise ValueError(f"Invalid subvolume number : {head}") from None
        if v >= n or v < 0:
            raise ValueError(f"Unexpected subvolume number. Got: {v}")
        toks = _split(rest)
        if len(toks) != s:
            raise ValueError(f"Parse of subvolume line failed (subvolume {rest.strip()}) has wrong length")
        for i, t in enumerate(toks):
            try:
                c = int(t)
            except ValueError:
                raise ValueError("parse of species count number failed") from None
            if c < 0:
                raise ValueError("Invalid: species count < 0")
            x[i, v] = c
    if x.max(initial=0) > 0xFFFFFFFF:
        raise ValueError("species count above 2^32-1")
    return {"N": n, "S": s, "x": x.astype(np.uint32).ravel()}


def read_reactions(path):
    with open(path) as f:
        lines = f.read().split("\n")
    r = _header(lines[0], "reactions")
    if r < 1:
        raise ValueError("Invalid: r < 1")
    s = _header(lines[1], "species")
    if s < 1:
        raise ValueError("Invalid: s < 1")
    reac = np.zeros((s, r), dtype=np.int32)
    prod = np.zeros((s, r), dtype=np.int32)
    j = 0
    for line in lines[3:]:
        if not line.strip():
            continue
        toks = _split(line)
        if len(toks) != 2 * s + 1:
            raise ValueError(f"Parse of reactions line failed (reaction {j})")
        if j >= r:
            raise ValueError(f"more reaction lines than reactions: {r}")
        try:
            reac[:, j] = [int(t) for t in toks[:s]]
            prod[:, j] = [int(t) for t in toks[s + 1:]]
        except ValueError:
            raise ValueError(f"parse of stechiometric coeff. failed (reaction {j})") from None
        j += 1
    return {"R": r, "S": s, "reac": reac.ravel(), "prod": prod.ravel()}


def read_constants(path, R, S):
    """Constants are single precision in the reference (std::stof into float
    arrays, cpp_utils.hpp:37,43); they are rounded to float32 here too."""
    with open(path) as f:
        l1 = f.readline()
        l2 = f.readline()
    k = [float(t) for t in _split(l1)]
    d = [float(t) for t in _split(l2)]
    if len(k) < R or len(d) < S:
        raise ValueError("constants file has too few values")
    return (np.array(k[:R], dtype=np.float32).astype(np.float64),
            np.array(d[:S], dtype=np.float32).astype(np.float64))


def read_compartments(path, N):
    comp = np.zeros(N, dtype=np.int32)
    with open(path) as f:
        lines = [l for l in f.read().split("\n") if l.strip()]
    if len(lines) < N:
        raise ValueError("compartments file too short")
    for i in range(N):
        comp[i] = int(lines[i].split(":", 1)[1])
    return comp


def read_conf(path, base=None):
    """Returns dict(net, topo, x0, t_end, log_freq) with the network in the
    C-ABI layout (plis_network)."""
    base = base if base is not None else os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        entries = [l.rstrip("\n") for l in f if l.strip()]
    it = iter(entries)

    def take(key, msg):
        try:
            line = next(it)
        except StopIteration:
            raise ValueError(msg) from None
        k, _, v = line.partition("=")
        if k != key:
            raise ValueError(msg)
        return v

    def fp(v):
        p = os.path.join(base, v)
        if not os.path.exists(p):
            raise ValueError(f"file {v} not found")
        return p

    topo = read_topology(fp(take("topology", "need a topology file")))
    rx = read_reactions(fp(take("reactions", "need a reactions file")))
    st = read_state(fp(take("initialState", "need a state file")))
    t_end = float(take("endTime", "need an endtime value"))
    log_freq = float(take("logFreq", "need a logFreq value (0 = no logging)"))
    C = int(take("compartments", "need a compartments value"))
    if C < 1:
        raise ValueError("compartments count < 1")
    comp = None
    if C > 1:
        comp = read_compartments(fp(take("compartmentsFile", "need a compartments file")), topo["N"])
    ks, ds = [], []
    for i in range(C):
        k, d = read_constants(fp(take("constants", f"need a constants file ({i + 1})")), rx["R"], rx["S"])
        ks.append(k)
        ds.append(d)
    if st["N"] != topo["N"] or st["S"] != rx["S"]:
        raise ValueError("state, topology and reactions disagree on N or S")
    net = {"S": rx["S"], "R": rx["R"], "C": C, "reac": rx["reac"], "prod": rx["prod"],
           "k": np.concatenate(ks), "d": np.concatenate(ds)}
    if comp is not None:
        topo = dict(topo, comp=comp)
    return {"net": net, "topo": topo, "x0": st["x"], "t_end": t_end, "log_freq": log_freq}


def snapshot_text(x, S, N, t):
    """The reference's snapshot format (PLISCuda/src/driver.cu:435-447):
    subvolumes=N, species=S, time=<std::to_string(float t)>, blank line, then
    one line per subvolume with a trailing space after every count."""
    x = np.asarray(x).reshape(S, N)
    head = f"subvolumes={N}\nspecies={S}\ntime={float(np.float32(t)):.6f}\n\n"
    body = "".join("".join(f"{int(c)} " for c in x[:, v]) + "\n" for v in range(N))
    return head + body
