"""Writes tests/golden/known_answers.json.

Each case is a list A[i] = i of n int32 items, a removal list R, and the
expected survivors A'[0, n-k).  Cases 'paper_fig3' and 'paper_fig2' are the
worked examples of the Redzone paper (RedZone paper.pdf p.4 Fig. 3 and p.2
Fig. 2, letters mapped to their indices).  The other cases are derived by hand
from the paper's algorithm (p.4 Alg. 1): the derivation is kept next to each
case so that it can be re-checked without running any code.
"""
import json
import os

CASES = [
    {
        "name": "paper_fig3",
        "n": 8, "R": [6, 0, 1, 4], "expect": [5, 7, 2, 3],
        "note": "A=[A,B,C,D,a,b,c,d], R=[6,0,1,4] -> [b,d,C,D] (paper p.4 Fig. 3). s_Z=4, Z=[a,b,c,d]; "
                "tail removals 6,4 kill Z[2]=c and Z[0]=a. i=0: d=6 in tail, Z[0] dead -> nothing; "
                "i=1: A[0]=Z[1]=b; i=2: Z[2] dead -> dest orphan 1; i=3: d=4 in tail -> source orphan Z[3]=d; "
                "pair: A[1]=d.",
    },
    {
        "name": "paper_fig2",
        "n": 7, "R": [2, 4], "expect": [0, 1, 5, 3, 6],
        "note": "A=[A,B,C,D,E,a,b], R=[2,4] -> [A,B,a,D,b] (paper p.2 Fig. 2 top). s_Z=5, no tail removal: "
                "A[2]=Z[0]=a, A[4]=Z[1]=b.",
    },
    {
        "name": "two_orphan_pairs",
        "n": 10, "R": [4, 1, 7, 6], "expect": [0, 9, 2, 3, 8, 5],
        "note": "s_Z=6, Z=A[6..10). Tail removals 7,6 kill Z[1], Z[0]. i=0: d=4, Z[0] dead -> dest orphan 4; "
                "i=1: d=1, Z[1] dead -> dest orphan 1; i=2: d=7 in tail, Z[2]=8 alive -> source orphan 8; "
                "i=3: d=6 in tail, Z[3]=9 alive -> source orphan 9. Pairs in order: A[4]=8, A[1]=9. "
                "Reverse pairing would give A[4]=9, A[1]=8.",
    },
    {
        "name": "interleaved_orphans",
        "n": 12, "R": [2, 8, 5, 9, 7], "expect": [0, 1, 10, 3, 4, 11, 6],
        "note": "s_Z=7, Z=A[7..12). Tail removals 8,9,7 kill Z[1], Z[2], Z[0]. i=0: d=2, dead -> dest orphan 2; "
                "i=1: d=8 tail, dead -> nothing; i=2: d=5, dead -> dest orphan 5; i=3: d=9 tail, Z[3]=10 alive "
                "-> source orphan 10; i=4: d=7 tail, Z[4]=11 alive -> source orphan 11. Pairs: A[2]=10, A[5]=11.",
    },
    {
        "name": "all_in_tail",
        "n": 8, "R": [7, 5, 6, 4], "expect": [0, 1, 2, 3],
        "note": "R is exactly the tail: every pair has a tail destination and a dead source -> nothing moves.",
    },
    {
        "name": "no_tail_removal",
        "n": 8, "R": [3, 0], "expect": [7, 1, 2, 6, 4, 5],
        "note": "s_Z=6, no tail removal: A[3]=Z[0]=6, A[0]=Z[1]=7.",
    },
    {
        "name": "k_zero", "n": 5, "R": [], "expect": [0, 1, 2, 3, 4],
        "note": "k=0 is a no-op (redzone.hxx:651).",
    },
    {
        "name": "k_equals_n", "n": 3, "R": [2, 0, 1], "expect": [],
        "note": "k=n leaves nothing (redzone.hxx:651).",
    },
    {
        "name": "single_item_removed", "n": 1, "R": [0], "expect": [],
        "note": "n=1, k=1.",
    },
    {
        "name": "last_item_removed", "n": 6, "R": [5], "expect": [0, 1, 2, 3, 4],
        "note": "The only removal is the tail item itself: dead source, tail destination -> nothing.",
    },
    {
        "name": "first_item_removed", "n": 6, "R": [0], "expect": [5, 1, 2, 3, 4],
        "note": "A[0]=Z[0]=5.",
    },
]


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(out, "w") as f:
        json.dump({"source": "Redzone paper (ACM TOPC 2024) Fig. 2-3 and hand derivations from its Alg. 1",
                   "cases": CASES}, f, indent=1)
    print(f"wrote {out} ({len(CASES)} cases)")


if __name__ == "__main__":
    main()
