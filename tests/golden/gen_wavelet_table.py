"""Generate the predefined-wavelet step table from the reference's wavelet header.

The reference defines its 34 wavelets as compile-time step tuples
(``src/lifted-wavelets.hpp:12-1107``) whose coefficients are ``long double``
literals cast to the transform type (``src/lifted-common.hpp:1000-1032``).
This script reads that header *as text*, extracts every step (kind, updated
band, offset, literal text of each tap, scale factors), the step order of the
``steps`` tuple and the filter ``width``, and writes

* ``lifted_amd/csrc/wavelet_table.inc`` -- the same numbers in this project's
  own step-program format (see ``wavelet_defs.hpp``), literals copied verbatim
  so that ``static_cast<T>(literal)`` reproduces the reference bits, including
  the literal typos the reference carries (Bior2_6 tap ``-0.076171874L``,
  Bior4_4 ``0.226562L``, the ``L``-less Symlet4 tap, Bior3_1's two s-updates);
* ``tests/golden/wavelet_steps.json`` -- the parsed table, used by the CPU
  tests to cross-check the compiled table without the reference present.

Run once in a checkout that has the reference at /root/reference:
    python tests/golden/gen_wavelet_table.py
"""
import json
import os
import re
import sys

REF = os.environ.get("LIFTED_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def parse_enum(common_text):
    block = re.search(r"#define LIFTED_WAVELETS\s*\\(.*?)#endif", common_text, re.S).group(1)
    return [(m.group(1), int(m.group(2))) for m in re.finditer(r"X\((\w+),\s*(\d+)\)", block)]


def split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def parse_wavelets(text):
    classes = {}
    # class bodies: "class Name{" ... "width = N;"
    for m in re.finditer(r"class\s+(\w+)\s*\{(.*?)width\s*=\s*(\d+)\s*;", text, re.S):
        name, body, width = m.group(1), m.group(2), int(m.group(3))
        defs = {}
        for d in re.finditer(
            r"constexpr\s+static\s+auto\s+(\w+)\s*=\s*detail::(\w+)\s*(<[^>]*>)?\s*\((.*?)\)\s*;",
            body, re.S,
        ):
            var, func, targs, args = d.group(1), d.group(2), d.group(3) or "", d.group(4)
            args = split_args(" ".join(args.split()))
            defs[var] = (func, targs, args)
        order_m = re.search(r"steps\s*=\s*std::make_tuple\((.*?)\)\s*;", body, re.S)
        order = [a for a in split_args(" ".join(order_m.group(1).split()))] if order_m else []
        steps = []
        for var in order:
            func, targs, args = defs[var]
            if func in ("update_s", "update_d"):
                steps.append({"kind": "generic", "target": "s" if func == "update_s" else "d",
                              "offset": int(args[0]), "taps": args[1:]})
            elif func in ("unit_update_s", "unit_update_d"):
                sign = "+" if "UpdateOperation::add" in targs else "-"
                steps.append({"kind": "unit", "target": "s" if func == "unit_update_s" else "d",
                              "offset": int(args[0]), "sign": sign})
            elif func in ("repeat_update_s", "repeat_update_d"):
                n = int(re.search(r",\s*(\d+)\s*>", targs).group(1))
                steps.append({"kind": "repeat", "target": "s" if func == "repeat_update_s" else "d",
                              "offset": int(args[0]), "n": n, "value": args[1]})
            elif func == "ScaleStep":
                steps.append({"kind": "scale", "values": args})
            else:
                raise ValueError(f"unknown step factory {func} in {name}")
        classes[name] = {"width": width, "steps": steps}
    return classes


def cxx_literal(tok):
    tok = tok.strip()
    if not re.fullmatch(r"-?[0-9.]+(e-?\d+)?L?", tok):
        raise ValueError(f"unexpected literal {tok!r}")
    return tok


def emit(enum, classes):
    lines = [
        "// Generated by tests/golden/gen_wavelet_table.py from the reference's",
        "// src/lifted-wavelets.hpp (coefficient literals kept verbatim). Do not edit.",
        "// Format: LFX_WAVELET(enum_name, enum_value, width) followed by its steps and",
        "// LFX_END(enum_name). Steps: LFX_GENERIC(band, offset, taps...),",
        "// LFX_UNIT(band, offset, sign), LFX_REPEAT(band, offset, n, value),",
        "// LFX_SCALE(s_mul, d_div).",
        "",
    ]
    for name, val in enum:
        c = classes[name]
        lines.append(f"LFX_WAVELET({name}, {val}, {c['width']})")
        for st in c["steps"]:
            if st["kind"] == "generic":
                taps = ", ".join(cxx_literal(t) for t in st["taps"])
                lines.append(f"  LFX_GENERIC({st['target'].upper()}, {st['offset']}, {taps})")
            elif st["kind"] == "unit":
                lines.append(f"  LFX_UNIT({st['target'].upper()}, {st['offset']}, {st['sign']}1)")
            elif st["kind"] == "repeat":
                lines.append(f"  LFX_REPEAT({st['target'].upper()}, {st['offset']}, {st['n']}, "
                             f"{cxx_literal(st['value'])})")
            else:
                vals = [cxx_literal(v) for v in st["values"]]
                if len(vals) == 1:
                    vals = vals * 2
                lines.append(f"  LFX_SCALE({vals[0]}, {vals[1]})")
        lines.append(f"LFX_END({name})")
    return "\n".join(lines) + "\n"


def main():
    wv = open(os.path.join(REF, "src", "lifted-wavelets.hpp")).read()
    common = open(os.path.join(REF, "src", "lifted-common.hpp")).read()
    enum = parse_enum(common)
    classes = parse_wavelets(wv)
    missing = [n for n, _ in enum if n not in classes]
    if missing:
        sys.exit(f"wavelets missing from parse: {missing}")
    with open(os.path.join(REPO, "lifted_amd", "csrc", "wavelet_table.inc"), "w") as f:
        f.write(emit(enum, classes))
    table = [{"name": n, "id": v, **classes[n]} for n, v in enum]
    with open(os.path.join(HERE, "wavelet_steps.json"), "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {len(enum)} wavelets")


if __name__ == "__main__":
    main()
