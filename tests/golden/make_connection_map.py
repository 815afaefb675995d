"""Write full_face_copy_map.json: the ghost-to-source cell mapping of the
reference's full-face block connections, read from the case tables of
src/eilmer/bc/ghost_cell_effect/full_face_copy.d:142-962
(set_up_cell_mapping_phase0), as data.

Each entry gives, for a destination face, a source face and (3D) an
orientation, how each in-face index of the source cell is formed from the
destination ghost cell's in-face indices: [destination index name, reversed],
where reversed means n - x - 1 with n the source block's cell count along that
axis.  The depth rule (layer 1 is the source face's boundary cell, layer 2 the
next one inward) is the same for every case and is not repeated per entry.

Run in a container that holds the reference:  python make_connection_map.py
"""
import json
import os
import re

SRC = "/root/reference/src/eilmer/bc/ghost_cell_effect/full_face_copy.d"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full_face_copy_map.json")
FACES = ["north", "east", "south", "west", "top", "bottom"]
IN_FACE = {"north": "ik", "south": "ik", "east": "jk", "west": "jk", "top": "ij", "bottom": "ij"}


def parse_expr(e):
    e = e.strip()
    m = re.fullmatch(r"other_blk\.n[ijk]cell - ([ijk]) - 1", e)
    if m:
        return [m.group(1), True]
    if re.fullmatch(r"[ijk]", e):
        return [e, False]
    raise ValueError("unexpected source index expression %r" % e)


def cases(text, pattern):
    found = [(m.start(), m.group(1)) for m in re.finditer(pattern, text)]
    for n, (pos, name) in enumerate(found):
        end = found[n + 1][0] if n + 1 < len(found) else len(text)
        yield name, text[pos:end]


def main():
    text = open(SRC).read()
    i2 = text.index("if (blk.myConfig.dimensions == 2) {")
    i3 = text.index("// presume dimensions == 3")
    dest_pat = r"case Face\.(\w+):\s*\n\s*[ijk]_dest = this_blk"
    src_pat = r"case Face\.(\w+):\s*\n\s*[ijk]_src = "
    entries = []
    for f, blk in cases(text[i2:i3], dest_pat):
        for g, body in cases(blk, src_pat):
            asg = dict(re.findall(r"([ijk])_src = ([^;]+);", body))
            ax = IN_FACE[g][0]
            entries.append({"dims": 2, "face": FACES.index(f), "other_face": FACES.index(g), "orientation": 0,
                            "src": {ax: parse_expr(asg[ax])}})
    for f, blk in cases(text[i3:], dest_pat):
        for g, body in cases(blk, src_pat):
            for o in range(4):
                line = re.search(r"case %d: ([^\n]*)" % o, body).group(1)
                asg = dict(re.findall(r"([ijk])_src = ([^;]+);", line))
                entries.append({"dims": 3, "face": FACES.index(f), "other_face": FACES.index(g), "orientation": o,
                                "src": {ax: parse_expr(asg[ax]) for ax in IN_FACE[g]}})
    json.dump({"source": "src/eilmer/bc/ghost_cell_effect/full_face_copy.d:142-962",
               "in_face_axes": IN_FACE, "entries": entries}, open(OUT, "w"), indent=0)
    print("%d entries -> %s" % (len(entries), OUT))


if __name__ == "__main__":
    main()
