"""Regenerate sobol_matrices.inc from the reference's Sobol generator matrices.

The table is published data (Joe & Kuo 2008, new-joe-kuo-6.21201 direction
numbers) that orion stores as 1024 dimensions x 52 uint32 columns in
src/sampler/sobol.cpp:17-53267.  Sample indices on the rendering path stay
below 2^32, so only the first 32 columns of each dimension are kept here.

Usage: python gen_sobol_table.py /root/reference/src/sampler/sobol.cpp
"""
import re
import sys
from pathlib import Path

N_DIMS = 1024
SRC_COLS = 52
KEEP_COLS = 32


def main(src_path: str) -> None:
    text = Path(src_path).read_text()
    start = text.index("matrices[Matrices::num_dimensions * Matrices::size] =")
    body = text[text.index("{", start) + 1:text.index("};", start)]
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]+)U", body)]
    if len(vals) != N_DIMS * SRC_COLS:
        raise SystemExit(f"unexpected table size {len(vals)}")
    out = Path(__file__).with_name("sobol_matrices.inc")
    lines = ["// Generated by gen_sobol_table.py: Joe-Kuo Sobol generator matrices,",
             f"// {N_DIMS} dimensions x {KEEP_COLS} columns (column i = image of index bit i).",
             f"#define ORION_SOBOL_DIMS {N_DIMS}",
             f"#define ORION_SOBOL_COLS {KEEP_COLS}"]
    lines.append("ORION_SOBOL_STORAGE unsigned int orion_sobol_matrices"
                 f"[{N_DIMS * KEEP_COLS}] = {{")
    for d in range(N_DIMS):
        cols = vals[d * SRC_COLS:d * SRC_COLS + KEEP_COLS]
        lines.append("  " + ",".join(f"0x{c:08x}u" for c in cols) + ",")
    lines.append("};")
    out.write_text("\n".join(lines) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         "/root/reference/src/sampler/sobol.cpp")
