"""Writes the name lists the ABI tests compare against.

Reads the reference sources as text (nothing is imported or run):
  src/neurongpu_C.h          -> c_api_names.txt   (exported C functions)
  pythonlib/neurongpu.py     -> py_api_names.txt  (public def names)
Run in a checkout that has the reference at /root/reference:
  python tests/golden/gen_api_names.py
"""
import os
import re

REF = os.environ.get("NEURONGPU_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def c_names():
    txt = open(os.path.join(REF, "src", "neurongpu_C.h")).read()
    return sorted(set(re.findall(r"\b(NeuronGPU_\w+)\s*\(", txt)))


def py_names():
    txt = open(os.path.join(REF, "pythonlib", "neurongpu.py")).read()
    return sorted(set(n for n in re.findall(r"^def (\w+)\s*\(", txt, re.M)
                      if not n.startswith("_")))


if __name__ == "__main__":
    for fn, names in (("c_api_names.txt", c_names()),
                      ("py_api_names.txt", py_names())):
        with open(os.path.join(HERE, fn), "w") as f:
            f.write("\n".join(names) + "\n")
        print(fn, len(names))
