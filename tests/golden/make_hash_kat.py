"""Writes hash_kat.json: known answers of cellCoordToHash (updateCellEntries.comp:36-52) for a
spread of cells, each with its keys under the masks of the tables the tests use.  The hash
is evaluated here in plain Python integers from the shader's published steps, independent
of the oracle's C code:  h = (x * 73856093) ^ (y * 19349663) ^ (z * 83492791) mod 2^32, then
h ^= h >> 16; h *= 0x7feb352d; h ^= h >> 15; h *= 0x846ca68b; h ^= h >> 16 (all mod 2^32).
Cells with negative coordinates are given as their uint32 wrap, as the force pass makes
them (uvec3 of a negative int)."""
import json
import os

M = 0xFFFFFFFF


def cell_hash(x: int, y: int, z: int) -> int:
    h = ((x * 73856093) & M) ^ ((y * 19349663) & M) ^ ((z * 83492791) & M)
    h ^= h >> 16
    h = (h * 0x7FEB352D) & M
    h ^= h >> 15
    h = (h * 0x846CA68B) & M
    h ^= h >> 16
    return h


def main() -> None:
    cells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (3, 5, 7), (M, 0, 0), (0, M, 0), (0, 0, M),
             (M, M, M), (113, 56, 56), (28, 17, 17), (457, 114, 228), (1000, 2000, 3000)]
    out = [{"cell": list(c), "hash": cell_hash(*c),
            "key": {str(b): cell_hash(*c) & ((1 << b) - 1) for b in (14, 20, 21, 25)}}
           for c in cells]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
