"""Table of LDS cycles per load from an lds_bank_probe counter run.

    python -m pyflct_amd.tune.lds_bank_table <rocprofv3 counter csv> <probe stdout log>

The probe (lds_bank_probe.hip) prints one "pattern: <name>" line per launch,
in launch order; the counter CSV holds SQ_LDS_IDX_ACTIVE and
SQ_LDS_BANK_CONFLICT per dispatch.  Each launch makes 4096 x 256 probe loads
plus the same fill of the LDS array; the first pattern (conflict-free
ds_read_b32) is the reference for the fill and counts as one cycle per load.
"""
import collections
import csv
import sys

LOADS = 4096 * 256


def table(csv_path, log_path):
    agg = collections.OrderedDict()
    with open(csv_path) as fh:
        for r in csv.DictReader(fh):
            if "k_probe" not in r["Kernel_Name"]:
                continue
            agg.setdefault(int(r["Dispatch_Id"]), collections.Counter())[r["Counter_Name"]] += \
                float(r["Counter_Value"])
    with open(log_path) as fh:
        names = [ln.split("pattern: ", 1)[1].strip() for ln in fh if ln.startswith("pattern: ")]
    if len(names) != len(agg):
        raise SystemExit(f"{len(names)} patterns in the log, {len(agg)} probe dispatches in the csv")
    rows = []
    base = None
    for (_, v), name in zip(sorted(agg.items()), names):
        act, bc = v["SQ_LDS_IDX_ACTIVE"], v["SQ_LDS_BANK_CONFLICT"]
        if base is None:
            base = (act, bc)
        rows.append((name, (act - base[0]) / LOADS + 1.0, (bc - base[1]) / LOADS))
    return rows


def main(argv):
    if len(argv) != 3:
        raise SystemExit(__doc__)
    print("| pattern | LDS active cycles per load | bank-conflict cycles per load |")
    print("|---|---|---|")
    for name, act, bc in table(argv[1], argv[2]):
        print(f"| {name} | {act:.2f} | {bc:.2f} |")


if __name__ == "__main__":
    main(sys.argv)
