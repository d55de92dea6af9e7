"""Extract the reference's own test vectors into committed fixtures.

Reads two data-bearing source files of texpresso 2.0.2 as text and writes
plain data (no reference code is copied):

  * lib/src/test_data.rs  -> kat.json: the 8 single-block known-answer sets
    (decoded RGBA pixels, encoded block) used by lib/src/lib.rs:363-504, plus
    the 4x6 height-not-multiple-of-4 decode case (lib/src/lib.rs:415-444) and
    the storage-size table (lib/src/lib.rs:349-361).
  * lib/src/colourfit/single_lut.rs -> single_lut.bin: the four 256-entry
    single-colour lookup tables, 6 bytes per entry
    (start0, end0, err0, start1, end1, err1), in the order
    LOOKUP_5_3, LOOKUP_6_3, LOOKUP_5_4, LOOKUP_6_4.

Run from the repository root:  python tests/golden/make_fixtures.py [REF_ROOT]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _array_after(text, marker):
    """Return the integers of the first [...] literal after `marker`."""
    pos = text.index(marker)
    if marker.startswith("const "):
        pos = text.index("= [", pos)
    start = text.index("[", pos)
    depth = 0
    for k in range(start, len(text)):
        if text[k] == "[":
            depth += 1
        elif text[k] == "]":
            depth -= 1
            if depth == 0:
                body = text[start + 1 : k]
                break
    body = re.sub(r"//[^\n]*", "", body)
    return [int(t, 0) for t in re.findall(r"0x[0-9A-Fa-f]+|\b\d+\b", body)]


def _encoded(text, name):
    seg = text[text.index("pub const %s:" % name) :]
    return _array_after(seg, "encoded:")


def _expand(luma):
    out = []
    for v in luma:
        out += [v, v, v]
    return out


def _add_alpha(rgb, alpha):
    out = []
    for i in range(16):
        out += rgb[3 * i : 3 * i + 3] + [alpha[i]]
    return out


def make_kat():
    text = open(os.path.join(REF, "lib/src/test_data.rs")).read()
    colour_rgb = _array_after(text, "const COLOUR_BLOCK_RGB")
    gray_luma = _array_after(text, "const GRAY_BLOCK_LUMA")
    ramp = _array_after(text, "const LINEAR_RAMP")
    bc3_alpha = _array_after(text, "const BC3_ALPHA_DECODED")
    opaque = [0xFF] * 16

    def inline_luma(name):
        seg = text[text.index("pub const %s:" % name) :]
        seg = seg[seg.index("decoded:") :]
        return _array_after(seg, "expand_single_to_rgb(&")

    def inline_rgb(name):
        seg = text[text.index("pub const %s:" % name) :]
        seg = seg[seg.index("decoded:") :]
        return _array_after(seg, "add_alpha_to_rgb(")

    sets = {
        "BC1_GRAY": (0, _add_alpha(_expand(inline_luma("BC1_GRAY")), opaque)),
        "BC1_COLOUR": (0, _add_alpha(colour_rgb, opaque)),
        "BC2_GRAY": (1, _add_alpha(_expand(gray_luma), ramp)),
        "BC2_COLOUR": (1, _add_alpha(colour_rgb, ramp)),
        "BC3_GRAY": (2, _add_alpha(_expand(gray_luma), bc3_alpha)),
        "BC3_COLOUR": (2, _add_alpha(colour_rgb, bc3_alpha)),
        "BC4_GRAY": (3, _add_alpha(_expand(inline_luma("BC4_GRAY")), opaque)),
        "BC5_GRAY": (4, _add_alpha(inline_rgb("BC5_GRAY"), opaque)),
    }
    kat = []
    for name, (fmt, decoded) in sets.items():
        enc = _encoded(text, name)
        assert len(decoded) == 64, name
        assert len(enc) in (8, 16), name
        kat.append({"name": name, "format": fmt, "decoded": decoded, "encoded": enc})

    lib = open(os.path.join(REF, "lib/src/lib.rs")).read()
    h4 = _array_after(lib, "let encoded =")
    sizes = []
    for m in re.finditer(r"Format::Bc(\d)\.compressed_size\((\d+), (\d+)\), (\d+)\)", lib):
        sizes.append([int(m.group(1)) - 1, int(m.group(2)), int(m.group(3)), int(m.group(4))])
    return {
        "source": "texpresso 2.0.2 lib/src/test_data.rs, lib/src/lib.rs:349-444",
        "blocks": kat,
        "decode_height6": {"format": 0, "width": 4, "height": 6, "encoded": h4, "pixel": [0x7F, 0x7F, 0x7F, 0xFF]},
        "compressed_size": sizes,
    }


def make_lut():
    text = open(os.path.join(REF, "lib/src/colourfit/single_lut.rs")).read()
    out = bytearray()
    for name in ("LOOKUP_5_3", "LOOKUP_6_3", "LOOKUP_5_4", "LOOKUP_6_4"):
        seg = text[text.index("pub const %s:" % name) :]
        seg = seg[seg.index("= [") :]
        end = seg.index("];")
        entries = re.findall(r"entry!\(([^)]*)\)", seg[:end])
        assert len(entries) == 256, name
        for e in entries:
            vals = [int(v) for v in e.split(",")]
            assert len(vals) == 6
            out += bytes(vals)
    return bytes(out)


if __name__ == "__main__":
    kat = make_kat()
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    with open(os.path.join(HERE, "single_lut.bin"), "wb") as f:
        f.write(make_lut())
    print("wrote kat.json (%d blocks) and single_lut.bin" % len(kat["blocks"]))
