"""Writes the DNG NoiseProfile fixtures under tests/golden/dng/ (run from the repo
root: python tests/golden/make_dng.py).

Each file is a minimal DNG: classic TIFF, a 16x16 16-bit CFA raw strip (a ramp),
DNGVersion 1.4, and NoiseProfile (tag 51041, DOUBLE [S0, O0, S1, O1, ...]) where
the case wants one. expected.json holds, per file, the pairs a reader must
return as (beta, lambda) = (O, S), or null for a file without the tag.
  le_ifd0.dng        little endian, raw data in IFD0, three per-plane pairs
  be_subifd.dng      big endian, IFD0 a preview (NewSubFileType 1) carrying a
                     decoy NoiseProfile, the raw image in a SubIFD (NewSubFileType
                     0) with one pair: the raw IFD's pairs win
  le_chain.dng       little endian, IFD0 a preview without the tag, the raw image
                     in the next IFD of the chain with four pairs (2x2 CFA planes)
  none.dng           little endian, no NoiseProfile
"""
import json
import os
import struct

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dng")
W = H = 16
SHORT, LONG, RATIONAL, DOUBLE, BYTE = 3, 4, 5, 12, 1


def raw_strip():
    return b"".join(struct.pack("<H", (x * 97 + y * 211) % 4096) for y in range(H) for x in range(W))


class Tiff:
    """Lays out a TIFF: IFDs are placed in order; values longer than 4 bytes go
    to a data area after all IFDs."""

    def __init__(self, big):
        self.e = ">" if big else "<"
        self.big = big

    def build(self, ifds):
        # ifds: list of (entries, kind) where entries = [(tag, type, values)], kind names
        # the link: ("chain" -> next IFD is the following one, "sub" -> referenced by SubIFDs)
        e = self.e
        sizes = [2 + 12 * len(ents) + 4 for ents, _ in ifds]
        offs = []
        pos = 8
        for sz in sizes:
            offs.append(pos)
            pos += sz
        data = bytearray()
        data_base = pos
        strip_off = None
        out = bytearray(b"MM" if self.big else b"II")
        out += struct.pack(e + "HI", 42, offs[0])
        blobs = []
        for k, (ents, link) in enumerate(ifds):
            ifd = bytearray(struct.pack(e + "H", len(ents)))
            ents = [(330, LONG, [offs[j] for j in v]) if t == "SUBIFD" else (t, ty, v) for t, ty, v in ents]
            for tag, typ, vals in sorted(ents, key=lambda t: t[0]):
                fmt = {SHORT: "H", LONG: "I", DOUBLE: "d", BYTE: "B"}[typ]
                if tag == 273:  # StripOffsets: patched to the strip position below
                    vals = [0]
                payload = struct.pack(e + fmt * len(vals), *vals)
                if len(payload) <= 4:
                    field = payload + b"\0" * (4 - len(payload))
                else:
                    field = struct.pack(e + "I", data_base + len(data))
                    data += payload
                    if len(data) % 2:
                        data += b"\0"
                if tag == 273:
                    blobs.append(len(out) + len(ifd) + 8)
                ifd += struct.pack(e + "HHI", tag, typ, len(vals)) + field
            nxt = offs[k + 1] if link == "chain" and k + 1 < len(ifds) else 0
            ifd += struct.pack(e + "I", nxt)
            out += ifd
        out += data
        strip_off = len(out)
        out += raw_strip()
        for at in blobs:
            out[at:at + 4] = struct.pack(e + "I", strip_off)
        return bytes(out)


def raw_entries(subfile, noise=None):
    ents = [(254, LONG, [subfile]), (256, LONG, [W]), (257, LONG, [H]), (258, SHORT, [16]), (259, SHORT, [1]),
            (262, SHORT, [32803]), (273, LONG, [0]), (277, SHORT, [1]), (279, LONG, [W * H * 2]),
            (50706, BYTE, [1, 4, 0, 0])]
    if noise is not None:
        flat = [v for s_o in noise for v in s_o]
        ents.append((51041, DOUBLE, flat))
    return ents


def main():
    os.makedirs(HERE, exist_ok=True)
    expected = {}
    # (S, O) pairs as the file stores them
    cases = {
        "le_ifd0.dng": (False, [(raw_entries(0, [(0.0123, 4.1e-6), (0.0098, 3.7e-6), (0.0151, 5.3e-6)]), "chain")],
                        [(0.0123, 4.1e-6), (0.0098, 3.7e-6), (0.0151, 5.3e-6)]),
        "be_subifd.dng": (True, [(raw_entries(1, [(0.5, 0.25)]) + [("SUBIFD", LONG, [1])], "none"),
                                 (raw_entries(0, [(0.0075, 2.5e-5)]), "sub")],
                          [(0.0075, 2.5e-5)]),
        "le_chain.dng": (False, [(raw_entries(1), "chain"),
                                 (raw_entries(0, [(0.011, 3e-6), (0.012, 3.5e-6), (0.010, 2.5e-6), (0.013, 4e-6)]),
                                  "chain")],
                         [(0.011, 3e-6), (0.012, 3.5e-6), (0.010, 2.5e-6), (0.013, 4e-6)]),
        "none.dng": (False, [(raw_entries(0), "chain")], None),
    }
    for name, (big, ifds, pairs) in cases.items():
        data = Tiff(big).build(ifds)
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(data)
        expected[name] = None if pairs is None else [[o, s] for s, o in pairs]
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
