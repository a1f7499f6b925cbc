"""Extract the KAT-1 known-answer data (95-halo dPIE model, one point, expected gradient) from the
reference benchmark source Benchmarks/GradientBenchmark/setup.cpp:19-44 into kat1_gradient.json.

Only the numeric arrays are read (this is a data fixture); run once in a container that has the
reference mounted:  python tests/golden/make_kat1.py /root/reference
"""
import json
import os
import re
import sys


def main(ref_root: str):
    src = open(os.path.join(ref_root, "Benchmarks/GradientBenchmark/setup.cpp")).read()

    def scalar(name):
        m = re.search(r"\b%s\s*=\s*([-+0-9.eE]+)\s*;" % name, src)
        return float(m.group(1))

    def array(name):
        # first (uncommented) definition of `name[] = { ... };`
        for m in re.finditer(r"^(?!\s*//).*?\b%s\s*\[\]\s*=\s*\{([^}]*)\}" % name, src, re.M):
            return [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
        raise KeyError(name)

    data = {
        "source": "Benchmarks/GradientBenchmark/setup.cpp:19-44",
        "x": scalar("x_c"), "y": scalar("y_c"),
        "type": [int(v) for v in array("lense_type")],
        "pos_x": array("pos_x"), "pos_y": array("pos_y"), "theta": array("theta"),
        "epot": array("epot"), "rcore": array("rcore"), "rcut": array("rcut"), "b0": array("b0"),
        "grad_x": scalar("grad_x"), "grad_y": scalar("grad_y"),
    }
    n = int(scalar("num_lenses"))
    for k in ("type", "pos_x", "pos_y", "theta", "epot", "rcore", "rcut", "b0"):
        assert len(data[k]) == n, (k, len(data[k]), n)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat1_gradient.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out, n, "halos")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
