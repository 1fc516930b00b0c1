"""Regenerate metal_presets.inc: RGB eta / k of orion's named metals.

orion converts tabulated spectral n, k data (luxRender / Sopra tables,
material/metal.cpp:19-71) to RGB with Spectrum::fromSampled and the CIE 1931
tables (core/spectrum.h:101-122, core/spectrum.cpp:61-809) when a material
names `metalName` (material/metal.cpp:74-140).  The resulting RGB triples are
what the renderer consumes; they are computed here by calling the reference's
own compiled conversion (oracle/_ref, built by `make -C oracle ref`) so the
values are the reference's bit for bit, and stored as a small table.

Usage: python gen_metal_presets.py
"""
import ctypes as C
from pathlib import Path

import numpy as np

NAMES = ["gold", "silver", "copper", "aluminium", "amorphousCarbon"]


def main() -> None:
    root = Path(__file__).resolve().parents[2]
    lib = C.CDLL(str(root / "oracle" / "_ref" / "libref_orion.so"))
    lib.ref_metal_preset.restype = C.c_int
    lib.ref_metal_preset.argtypes = [C.c_char_p, np.ctypeslib.ndpointer(np.float32)]
    lines = ["// Generated by gen_metal_presets.py: RGB eta and k of orion's named metals",
             "// (Spectrum::fromSampled of the tabulated spectra), float bit patterns.",
             "struct OrionMetalPreset { const char *name; unsigned eta[3]; unsigned k[3]; };",
             "static const OrionMetalPreset orion_metal_presets[] = {"]
    for n in NAMES:
        out = np.zeros(6, np.float32)
        if lib.ref_metal_preset(n.encode(), out) != 0:
            raise SystemExit(f"preset {n} failed")
        bits = out.view(np.uint32)
        vals = ", ".join(f"{v:.9g}f" for v in out)
        lines.append(f"  {{\"{n}\", {{0x{bits[0]:08x}u, 0x{bits[1]:08x}u, 0x{bits[2]:08x}u}}, "
                     f"{{0x{bits[3]:08x}u, 0x{bits[4]:08x}u, 0x{bits[5]:08x}u}}}},  // {vals}")
    lines.append("};")
    out_path = Path(__file__).with_name("metal_presets.inc")
    out_path.write_text("\n".join(lines) + "\n")
    print(f"wrote {out_path}")


if __name__ == "__main__":
    main()
