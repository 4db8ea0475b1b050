"""Writes hand_table_ref.txt: the 4824 integers of the reference's sorted
7-card encoding table (crates/range-solver/src/hand_table.rs:1), one per line.

Run in a container where the reference is mounted; the fixture it writes is
committed, so tests never read the reference themselves.
"""
import re
import sys

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/crates/range-solver/src/hand_table.rs"
text = open(src).read()
body = text[text.index("["):]
body = body[body.index("= [") + 3:body.rindex("]")]
vals = [int(v) for v in re.findall(r"-?\d+", body)]
assert len(vals) == 4824, len(vals)
out = __file__.replace("make_hand_table_fixture.py", "hand_table_ref.txt")
with open(out, "w") as f:
    f.write("\n".join(str(v) for v in vals) + "\n")
print(out, len(vals))
