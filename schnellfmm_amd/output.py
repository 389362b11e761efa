"""Output stage: the calibrated results file written after the association
pass (OutputConfig / OutputFormat, src/testing.rs:33-74; coalesce_output_tsv
:740-801; coalesce_output_parquet :818-926).

TSV is written by the library (slfmm_write_tsv: several host threads format
rows in Rust's `{:.6e}` form).  Parquet uses pyarrow with the reference's
calibrated schema (src/testing.rs:621-640), Zstd compression and its row
order: worker w of W = min(n_workers, n_chunks) (at least 1) encodes chunks
w, w + W, ... of chunk_size SNPs, and the worker files are concatenated in
worker order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence, Union

import numpy as np

from ._lib import lib

TSV = "tsv"
PARQUET = "parquet"


def parse_output_format(s: str) -> str:
    """OutputFormat::from_str (src/testing.rs:51-59)."""
    v = s.lower()
    if v not in (TSV, PARQUET):
        raise ValueError(f"unknown output format '{s}', expected 'tsv' or 'parquet'")
    return v


@dataclass
class BimRecords:
    """The .bim fields the results carry (BimRecord, src/bed.rs:29-36)."""
    chrom: Sequence[str]
    pos: np.ndarray
    snp_id: Sequence[str]

    @classmethod
    def read(cls, bim_path: str) -> "BimRecords":
        """parse_bim (src/bed.rs:480-503) for the three output fields."""
        chrom, pos, ids = [], [], []
        with open(bim_path) as f:
            for i, line in enumerate(f):
                fld = line.split()
                if len(fld) < 6:
                    raise ValueError(f"BIM line {i + 1} has {len(fld)} fields, expected 6")
                chrom.append(fld[0])
                ids.append(fld[1])
                pos.append(int(fld[3]))
        return cls(chrom, np.asarray(pos, np.uint64), ids)


def _bim_path_for(bed_or_bim: str) -> str:
    if bed_or_bim.endswith(".bed"):
        return bed_or_bim[:-4] + ".bim"
    return bed_or_bim


def parquet_row_order(p: int, chunk_size: int, n_workers: int) -> np.ndarray:
    n_chunks = -(-p // chunk_size) if p else 0
    w_count = max(min(n_workers, n_chunks), 1)
    parts = []
    for w in range(w_count):
        for seq in range(w, n_chunks, w_count):
            parts.append(np.arange(seq * chunk_size, min(p, (seq + 1) * chunk_size), dtype=np.int64))
    return np.concatenate(parts) if parts else np.zeros(0, np.int64)


def write_results(path: str, results, bim: Union[str, BimRecords], cov_names: Optional[Sequence[str]] = None,
                  fmt: str = TSV, n_workers: int = 0, chunk_size: int = 10_000, threads: int = 0) -> None:
    """Write the calibrated results of a fit (TestResults) to `path`.

    bim: a .bim (or its .bed) path, or BimRecords.  cov_names default to
    cov_0, cov_1, ...  fmt: "tsv" or "parquet".  n_workers / chunk_size set
    the Parquet row order as the reference's Lfmm2Config fields do."""
    fmt = parse_output_format(fmt)
    tstat = np.ascontiguousarray(results.tstat, dtype=np.float64)
    p, d = tstat.shape
    pv = np.ascontiguousarray(results.pvalue, dtype=np.float64)
    be = np.ascontiguousarray(results.beta, dtype=np.float64)
    r2 = np.ascontiguousarray(results.r2, dtype=np.float64)
    names = list(cov_names) if cov_names is not None else [f"cov_{j}" for j in range(d)]
    if len(names) != d:
        raise ValueError(f"{len(names)} covariate names for {d} covariates")
    if fmt == TSV:
        cn = (C.c_char_p * d)(*[s.encode() for s in names])
        if isinstance(bim, str):
            rc = lib().slfmm_write_tsv(None, path.encode(), _bim_path_for(bim).encode(), None, None, None, p, d, cn,
                                       pv.ctypes.data, be.ctypes.data, tstat.ctypes.data, r2.ctypes.data, 0, threads)
        else:
            if len(bim.chrom) != p or len(bim.snp_id) != p or len(bim.pos) != p:
                raise ValueError("bim records do not match the number of SNPs")
            ch = (C.c_char_p * p)(*[s.encode() for s in bim.chrom])
            ids = (C.c_char_p * p)(*[s.encode() for s in bim.snp_id])
            pos = np.ascontiguousarray(bim.pos, dtype=np.uint64)
            rc = lib().slfmm_write_tsv(None, path.encode(), None, ch, pos.ctypes.data, ids, p, d, cn,
                                       pv.ctypes.data, be.ctypes.data, tstat.ctypes.data, r2.ctypes.data, 0, threads)
        if rc != 0:
            from ._lib import SlfmmError
            raise SlfmmError(rc, lib().slfmm_last_error(None).decode())
        return
    import pyarrow as pa
    import pyarrow.parquet as pq
    recs = BimRecords.read(_bim_path_for(bim)) if isinstance(bim, str) else bim
    if len(recs.chrom) != p:
        raise ValueError("bim records do not match the number of SNPs")
    order = parquet_row_order(p, chunk_size, n_workers)
    cols = {"chr": pa.array(np.asarray(recs.chrom, dtype=object)[order], pa.string()),
            "pos": pa.array(np.asarray(recs.pos).astype(np.int64)[order], pa.int64()),
            "snp_id": pa.array(np.asarray(recs.snp_id, dtype=object)[order], pa.string())}
    for j, nm in enumerate(names):
        cols[f"p_{nm}"] = pa.array(pv[order, j])
    for j, nm in enumerate(names):
        cols[f"beta_{nm}"] = pa.array(be[order, j])
    for j, nm in enumerate(names):
        cols[f"t_{nm}"] = pa.array(tstat[order, j])
    for k, nm in enumerate(("r2_cov", "r2_latent", "r2_resid")):
        cols[nm] = pa.array(r2[order, k])
    schema = pa.schema([pa.field(k, v.type, nullable=False) for k, v in cols.items()])
    table = pa.Table.from_arrays(list(cols.values()), schema=schema)
    pq.write_table(table, path, compression="zstd")
