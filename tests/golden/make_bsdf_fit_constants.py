"""Writes bsdf_fit_constants.json: the numeric literals of the reference's
fitted BSDF functions, in source order, as data.

The fits have no closed form to test against (Zeltner et al. 2022's sheen
albedo and LTC coefficients, OpenPBR Eq. 102's average-Fresnel fit), so the
oracle's and the device's constants are compared with these numbers, read
from the reference's GLSL:

  src/material/fuzz_brdf.glsl     fuzz_albedo, ltc_inverse_coeffs
  src/material/fresnel.glsl       dieletric_fresnel_factor
  src/material/diffuse_brdf.glsl  oren_nayar, energy_compensated_oren_nayar
  src/material/metal_brdf.glsl    fresnel_metal

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_bsdf_fit_constants.py
"""
import json
import os
import re

REFERENCE = "/root/reference/src/material"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bsdf_fit_constants.json")
FUNCTIONS = {
    "fuzz_brdf.glsl": ["fuzz_albedo", "ltc_inverse_coeffs"],
    "fresnel.glsl": ["dieletric_fresnel_factor"],
    "diffuse_brdf.glsl": ["oren_nayar", "energy_compensated_oren_nayar"],
    "metal_brdf.glsl": ["fresnel_metal"],
}
# Unsigned decimal literals not glued to an identifier (so pow2, vec3 and
# the like are skipped); a sign is a separate operator in both languages.
LITERAL = re.compile(r"(?<![\w.])(\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?(?![\w.])")


def function_body(text, name):
    m = re.search(r"^\w+\s+" + name + r"\s*\(", text, re.M)
    start = text.index("{", m.end())
    depth, i = 1, start + 1
    while depth:
        depth += {"{": 1, "}": -1}.get(text[i], 0)
        i += 1
    first_line = text.count("\n", 0, m.start()) + 1
    last_line = text.count("\n", 0, i) + 1
    return text[start:i], first_line, last_line


def literals(body):
    body = re.sub(r"//[^\n]*", "", body)
    return [float(m.group(0)) for m in LITERAL.finditer(body)]


def main():
    out = {}
    for file, names in FUNCTIONS.items():
        text = open(os.path.join(REFERENCE, file)).read()
        for name in names:
            body, a, b = function_body(text, name)
            out[name] = {"source": f"src/material/{file}:{a}-{b}", "literals": literals(body)}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
