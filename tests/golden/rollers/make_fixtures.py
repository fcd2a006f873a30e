"""Writes the rollers-example fixtures of this folder from the reference's
own data files (multi_bodies/examples/rollers/inputfile_rollers.dat,
multi_bodies/Structures/blob.vertex and the first 256 bodies of
multi_bodies/examples/rollers/rollers.clones, with the count line set to
256).  Run once where the reference tree is present:

    python tests/golden/rollers/make_fixtures.py /root/reference
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    ex = os.path.join(ref_root, "multi_bodies", "examples", "rollers")
    shutil.copyfile(os.path.join(ex, "inputfile_rollers.dat"), os.path.join(HERE, "inputfile_rollers.dat"))
    shutil.copyfile(os.path.join(ref_root, "multi_bodies", "Structures", "blob.vertex"),
                    os.path.join(HERE, "blob.vertex"))
    with open(os.path.join(ex, "rollers.clones")) as f:
        lines = f.read().splitlines()
    with open(os.path.join(HERE, "rollers_256.clones"), "w") as f:
        f.write("256\n")
        f.write("\n".join(lines[1:257]) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
