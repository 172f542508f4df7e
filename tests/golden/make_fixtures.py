"""Extract the reference's known-answer data into tests/golden/reference_fixtures.json.

The reference test (tests/test_bri17.cpp) holds four Maxima-generated element
matrices as comma-separated literals after ``Ke <<`` / ``Be <<``:

    2D Ke  8x8    tests/test_bri17.cpp:344-356
    3D Ke 24x24   tests/test_bri17.cpp:372-531
    2D Be  3x8    tests/test_bri17.cpp:547-552
    3D Be 6x24    tests/test_bri17.cpp:570-598

and the parameters they were generated with (mu=5.6, nu=0.3, shapes (3,4) and
(3,4,5), cell spacing (1.1,1.2,1.3), :319-333) and the tolerances of the
comparison (rtol 1e-15, atol 1e-14, :360,535,558,604; |imag| <= 1e-14, :140).

Only the numbers are written out (data, not source).  Run in a container that
has the reference mounted:  python tests/golden/make_fixtures.py [REF_ROOT]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _blocks(text, name):
    out = []
    for m in re.finditer(r"\b%s\s*<<(.*?);" % name, text, flags=re.S):
        body = re.sub(r"//[^\n]*", "", m.group(1))
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        out.append(vals)
    return out


def main(ref_root="/root/reference"):
    with open(os.path.join(ref_root, "tests", "test_bri17.cpp")) as f:
        text = f.read()
    ke = _blocks(text, "Ke")
    be = _blocks(text, "Be")
    assert [len(b) for b in ke] == [64, 576], [len(b) for b in ke]
    assert [len(b) for b in be] == [24, 144], [len(b) for b in be]
    data = {
        "source": "sbrisard/bri17 v0.1 tests/test_bri17.cpp (Maxima output of "
                  "reference/Q4_Q8_element.mac)",
        "mu": 5.6,
        "nu": 0.3,
        "shape2": [3, 4],
        "spacing2": [1.1, 1.2],
        "shape3": [3, 4, 5],
        "spacing3": [1.1, 1.2, 1.3],
        "rtol": 1e-15,
        "atol": 1e-14,
        "imag_tol": 1e-14,
        "Ke2": ke[0],
        "Ke3": ke[1],
        "Be2": be[0],
        "Be3": be[1],
    }
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as f:
        json.dump(data, f, indent=0)
        f.write("\n")
    # the same matrices as plain text (row-major, one value per line) for the
    # C++ drop-in test, which has no JSON parser
    for name in ("Ke2", "Ke3", "Be2", "Be3"):
        with open(os.path.join(HERE, name + ".txt"), "w") as f:
            f.write("\n".join(repr(v) for v in data[name]) + "\n")


if __name__ == "__main__":
    main(*sys.argv[1:])
