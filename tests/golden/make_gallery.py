"""Record the reference's gallery: command lines and image digests.

The reference ships 32 rendered stills under gallery/ together with the
command lines that made them (Makefile:33-110, targets `images` and
`stars`).  This script reads both from /root/reference and writes
tests/golden/gallery.json: per image the program arguments (without --out)
and the PNG file's length and SHA-256.  The images themselves are the
fixture's data; only their digests are kept, since the GPU drop-in program
reproduces the files byte for byte.
"""
import hashlib
import json
import os
import shlex

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def gallery_commands(makefile):
    """(file name, argument list) of every `./$(BIN) ... --out gallery/X`."""
    text = open(makefile).read().replace("\\\n", "")
    out = []
    for line in text.split("\n"):
        line = line.strip()
        if not line.startswith("./$(BIN)") or "--out gallery/" not in line:
            continue
        args = shlex.split(line)[1:]
        i = args.index("--out")
        name = os.path.basename(args[i + 1])
        del args[i:i + 2]
        out.append((name, args))
    return out


def main():
    cases = []
    for name, args in gallery_commands(os.path.join(REF, "Makefile")):
        data = open(os.path.join(REF, "gallery", name), "rb").read()
        cases.append({"name": name, "args": args, "bytes": len(data),
                      "sha256": hashlib.sha256(data).hexdigest()})
    out = {"note": "written by tests/golden/make_gallery.py from the reference's Makefile "
                   "(targets images, stars) and gallery/*.png",
           "cases": cases}
    with open(os.path.join(HERE, "gallery.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(cases)} gallery images")


if __name__ == "__main__":
    main()
