"""Per-kernel summaries of rocprofv3 counter dumps.

HBM traffic: rocprofv3 ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` report
KiB per dispatch.  Each counter is scaled by what it reads for a plain copy
of known size (``bench.py --calibrate-copy MIB``).

Instruction mix: ``--pmc SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_WAVES``
report wave-level instruction counts per dispatch; the summary keeps their
per-kernel means, from which a consumer derives the VALU issue fraction
(VALU instructions / (duration x 1,024 SIMDs x 2.4 GHz / 4 cycles)).

Every summary is stamped with the SHA-256 of the HIP library that was
profiled, so a consumer can tell whether the numbers describe the library
it is running.

    python -m emsim_amd.profiling FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv \
        --copy-mib 256 --out traffic.json
    python -m emsim_amd.profiling --sq SQ.csv --out valu.json
"""
import argparse
import csv
import hashlib
import json
from collections import defaultdict
from pathlib import Path


def library_sha256(path=None) -> str:
    if path is None:
        from .backend import hip_ext
        path = hip_ext.LIB_PATH
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def _per_kernel(path, counter):
    sums, counts = defaultdict(float), defaultdict(int)
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] != counter:
            continue
        name = row["Kernel_Name"]
        sums[name] += float(row["Counter_Value"])
        counts[name] += 1
    return {k: (sums[k] / counts[k], counts[k]) for k in sums}


def copy_factor(path, counter, nbytes) -> float:
    """bytes moved / counter reading (KiB) for the calibration copy."""
    rows = _per_kernel(path, counter)
    name, (kib, _) = max(rows.items(), key=lambda kv: kv[1][0])
    return nbytes / (kib * 1024.0)


def summarize(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_mib, lib=None):
    nbytes = copy_mib * (1 << 20)
    ff = copy_factor(copy_fetch_csv, "FETCH_SIZE", nbytes)
    fw = copy_factor(copy_write_csv, "WRITE_SIZE", nbytes)
    fetch = _per_kernel(fetch_csv, "FETCH_SIZE")
    write = _per_kernel(write_csv, "WRITE_SIZE")
    kernels = {}
    for name in sorted(set(fetch) | set(write)):
        rd = fetch.get(name, (0.0, 0))[0] * 1024.0 * ff
        wr = write.get(name, (0.0, 0))[0] * 1024.0 * fw
        kernels[name] = {"read_bytes": rd, "write_bytes": wr, "bytes": rd + wr,
                         "dispatches": max(fetch.get(name, (0, 0))[1], write.get(name, (0, 0))[1])}
    return {"library_sha256": library_sha256(lib), "fetch_scale": ff, "write_scale": fw,
            "kernels": kernels}


SQ_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_WAVES")
# also kept when the run collected them: cycles the LDS stalled on bank
# conflicts, wave-cycles waiting to issue LDS instructions, wave-cycles
# resident, and busy cycles
SQ_OPTIONAL = ("SQ_LDS_BANK_CONFLICT", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES")

# VALU issue peak of an MI355X in wave64 instructions per second: 256 CUs x 4
# SIMDs, one wave64 VALU instruction per SIMD every 4 cycles at 2.4 GHz
# (packed FP32 is disabled in this library, so no instruction does more).
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4


def summarize_sq(sq_csv, lib=None):
    """Per-kernel mean of each SQ instruction counter per dispatch."""
    kernels = {}
    for counter in SQ_COUNTERS + SQ_OPTIONAL:
        for name, (mean, count) in _per_kernel(sq_csv, counter).items():
            rec = kernels.setdefault(name, {"dispatches": count})
            rec[counter] = mean
    return {"library_sha256": library_sha256(lib), "valu_issue_peak_per_s": VALU_ISSUE_PEAK,
            "kernels": kernels}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("csv", nargs="*")
    ap.add_argument("--sq", default=None, help="SQ instruction-counter CSV to summarise")
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--lib", default=None)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.sq:
        res = summarize_sq(a.sq, a.lib)
        Path(a.out).write_text(json.dumps(res, indent=1) + "\n")
        print(f"wrote {a.out}: {len(res['kernels'])} kernels")
        return
    if len(a.csv) != 4 or a.copy_mib <= 0:
        ap.error("traffic summary needs FETCH WRITE COPY_FETCH COPY_WRITE csv files and "
                 "--copy-mib")
    res = summarize(*a.csv, a.copy_mib, a.lib)
    Path(a.out).write_text(json.dumps(res, indent=1) + "\n")
    print(f"wrote {a.out}: {len(res['kernels'])} kernels, fetch x{res['fetch_scale']:.3f}, "
          f"write x{res['write_scale']:.3f}")


if __name__ == "__main__":
    main()
