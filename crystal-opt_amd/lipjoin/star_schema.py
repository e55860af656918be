"""Synthetic N-dimension star schemas as star-join queries.

The descriptor follows the reference's star_schema_t
(src/ops/utils/star.cuh:25-64): a fact table of n_fact rows and up to four
dimension tables, dimension d with dim_rows[d] rows and a selection bound
dim_sels[d] (star.cuh:32-35: the predicate is `t.value <= sel`), plus
bloom_factor (filter bits per dimension row) and htable_factor (int32 words
of hash table per dimension row; star.cuh:121-123 sizes the allocations by
them).  The columns are the reference's (star.cuh:67-120,
star_table_init_host): fact keys 1..n_fact shuffled and values 0..n_fact-1;
dimension keys 1..dim_rows[d] shuffled and values 0..dim_rows[d]-1; the fact's
foreign keys into dimension d in shuffled blocks of 1..dim_rows[d].

The reference's star.cu builds such a schema and stops there.  The query
run over it here is the star join the engine runs: SUM(fact value) and the
row count over the fact rows whose every foreign key joins a selected
dimension row; PHT_2 tables of dim_rows * htable_factor / 2 slots (keys_min
0, so key k goes to slot k mod ht_len) and modulo filters of
dim_rows * bloom_factor bits.

gen="reference" draws the columns with the reference's own generator
(lj_gen_reference_star_host: srand(0) and the C library's rand(), one host
thread); gen="seeded" uses the counter-based generator instead (dimension
keys shuffled by lj_gen_shuffle_host, foreign keys uniform over 1..rows),
which is multithreaded and depends on the seed only.
"""
import numpy as np

from . import _abi
from . import (CMP_LTE, COMB_SET, GEN_SEQUENCE, GEN_UNIFORM, LJError, MAX_DIMS, dim_desc, gen_host, lib,
               shuffle_host, star_desc)


class StarSchema:
    """n_fact fact rows, dim_rows[d] rows and selection bound dim_sels[d] per
    dimension (1 to 4 of them), bloom_factor filter bits and htable_factor
    hash-table words per dimension row."""

    def __init__(self, n_fact, dim_rows, dim_sels, bloom_factor=2, htable_factor=2, gen="reference", seed=0):
        if not 1 <= len(dim_rows) <= MAX_DIMS or len(dim_sels) != len(dim_rows):
            raise ValueError("1 to 4 dimensions, one selection bound each")
        if bloom_factor < 1 or htable_factor < 2 or htable_factor % 2:
            raise ValueError("bloom_factor >= 1, htable_factor an even number >= 2 (PHT_2: 2 words a slot)")
        if gen not in ("reference", "seeded"):
            raise ValueError(f"unknown generator {gen!r}")
        self.n_fact = int(n_fact)
        self.dim_rows = [int(r) for r in dim_rows]
        self.dim_sels = [int(s) for s in dim_sels]
        self.bloom_factor, self.htable_factor = int(bloom_factor), int(htable_factor)
        self.gen, self.seed = gen, int(seed)

    @property
    def n_dims(self):
        return len(self.dim_rows)

    def ht_len(self, d):
        return self.dim_rows[d] * self.htable_factor // 2

    def bloom_bits(self, d):
        return self.dim_rows[d] * self.bloom_factor

    def predicate(self, d):
        return ((COMB_SET, CMP_LTE, self.dim_sels[d]),)

    # -- columns ------------------------------------------------------------------------
    def host_columns(self):
        """dict: fact_keys, fact_vals (numpy int32), and per dimension lists
        dim_keys, dim_vals, fact_fks."""
        nd, n = self.n_dims, self.n_fact
        out = {"fact_keys": np.empty(n, np.int32), "fact_vals": np.empty(n, np.int32),
               "dim_keys": [np.empty(r, np.int32) for r in self.dim_rows],
               "dim_vals": [np.empty(r, np.int32) for r in self.dim_rows],
               "fact_fks": [np.empty(n, np.int32) for _ in self.dim_rows]}
        if self.gen == "reference":
            rows = (_abi.C.c_int64 * nd)(*self.dim_rows)
            arr = lambda xs: (_abi.C.c_void_p * nd)(*[x.ctypes.data if x.size else None for x in xs])
            st = lib().lj_gen_reference_star_host(
                n, nd, rows, out["fact_keys"].ctypes.data if n else None,
                out["fact_vals"].ctypes.data if n else None, arr(out["dim_keys"]), arr(out["dim_vals"]),
                arr(out["fact_fks"]))
            if st != _abi.OK:
                raise LJError(st, lib().lj_last_error(None).decode())
            return out
        # seeded: stream 0 the fact keys, 1 + 2d dimension d's keys, 2 + 2d its foreign keys
        out["fact_keys"][:] = gen_host(GEN_SEQUENCE, self.seed, 0, 1, 0, n)
        shuffle_host(out["fact_keys"], self.seed, 0)
        out["fact_vals"][:] = gen_host(GEN_SEQUENCE, self.seed, 0, 0, 0, n)
        for d, r in enumerate(self.dim_rows):
            out["dim_keys"][d][:] = gen_host(GEN_SEQUENCE, self.seed, 1 + 2 * d, 1, 0, r)
            shuffle_host(out["dim_keys"][d], self.seed, 1 + 2 * d)
            out["dim_vals"][d][:] = gen_host(GEN_SEQUENCE, self.seed, 1 + 2 * d, 0, 0, r)
            out["fact_fks"][d][:] = gen_host(GEN_UNIFORM, self.seed, 2 + 2 * d, 1, r, n)
        return out

    @staticmethod
    def to_device(cols, device="cuda"):
        """The same dict with torch int32 tensors on `device`."""
        import torch
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        return {k: ([t(a) for a in v] if isinstance(v, list) else t(v)) for k, v in cols.items()}

    # -- the query -----------------------------------------------------------------------
    def dim_descs(self, dev, method="lip"):
        """lj_dim_desc of each dimension over device columns `dev`: its values
        are the predicate's attribute and the payload."""
        return [dim_desc(dev["dim_keys"][d], self.dim_rows[d], self.ht_len(d), attr=dev["dim_vals"][d],
                         payload=dev["dim_vals"][d], predicate=self.predicate(d), keys_min=0,
                         bloom=("mod", self.bloom_bits(d)) if method != "naive" else None)
                for d in range(self.n_dims)]

    def star_desc(self, dev, method="lip", order=None, dims=None):
        """lj_star_desc: SUM(fact value) and COUNT over the rows that join a
        selected row of every dimension (one group), probed in `order`."""
        order = list(order) if order is not None else list(range(self.n_dims))
        dims = dims if dims is not None else self.dim_descs(dev, method)
        return star_desc(dev["fact_fks"], dev["fact_vals"], self.n_fact, dims, order, method=method)
