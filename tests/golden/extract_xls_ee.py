"""Extract the EE profiles and timing table recorded in the reference's spreadsheet.

Reads ``report/cva-bcn-accuracy-mean-squares_code_fixed (2).xls`` (Excel BIFF8 inside an
OLE2 compound file) with a minimal parser written for this purpose (xlrd is not
available), and writes ``xls_ee_profiles.json`` beside this script. Only numeric cells
and shared-string labels are decoded; formulas contribute their cached numeric result.

The output is data: the CPU / GPU expected-exposure rows the reference's author
measured (51 values each), and the Resume sheet's timing table.

Usage:  python tests/golden/extract_xls_ee.py [path-to-xls]
"""
import json
import struct
import sys
from pathlib import Path

DEFAULT_XLS = "/root/reference/report/cva-bcn-accuracy-mean-squares_code_fixed (2).xls"


def _ole_stream(data: bytes, want: str) -> bytes:
    """Return the named stream of an OLE2 compound file (regular-sector streams only)."""
    if data[:8] != bytes.fromhex("D0CF11E0A1B11AE1"):
        raise ValueError("not an OLE2 compound file")
    sec_shift = struct.unpack_from("<H", data, 30)[0]
    mini_shift = struct.unpack_from("<H", data, 32)[0]
    ssz = 1 << sec_shift
    n_fat = struct.unpack_from("<I", data, 44)[0]
    dir_start = struct.unpack_from("<I", data, 48)[0]
    mini_cutoff = struct.unpack_from("<I", data, 56)[0]
    minifat_start = struct.unpack_from("<I", data, 60)[0]
    difat_start = struct.unpack_from("<I", data, 68)[0]
    difat = list(struct.unpack_from("<109I", data, 76))
    nxt = difat_start
    while nxt < 0xFFFFFFFA:
        off = 512 + nxt * ssz
        vals = struct.unpack_from("<%dI" % (ssz // 4), data, off)
        difat.extend(vals[:-1])
        nxt = vals[-1]
    fat_secs = [s for s in difat if s < 0xFFFFFFFA][:n_fat]
    fat = []
    for s in fat_secs:
        fat.extend(struct.unpack_from("<%dI" % (ssz // 4), data, 512 + s * ssz))

    def chain(start):
        out, s = [], start
        while s < 0xFFFFFFFA:
            out.append(s)
            s = fat[s]
        return out

    def read_chain(start):
        return b"".join(data[512 + s * ssz: 512 + (s + 1) * ssz] for s in chain(start))

    dirs = read_chain(dir_start)
    root_start = None
    for i in range(len(dirs) // 128):
        e = dirs[i * 128:(i + 1) * 128]
        nlen = struct.unpack_from("<H", e, 64)[0]
        name = e[:max(nlen - 2, 0)].decode("utf-16-le")
        etype = e[66]
        start, size = struct.unpack_from("<II", e, 116)
        if etype == 5:
            root_start = start
        if name == want:
            if size < mini_cutoff:
                mini = read_chain(root_start)
                mfat = []
                mf = read_chain(minifat_start)
                mfat = list(struct.unpack_from("<%dI" % (len(mf) // 4), mf))
                msz = 1 << mini_shift
                out, s = [], start
                while s < 0xFFFFFFFA:
                    out.append(mini[s * msz:(s + 1) * msz])
                    s = mfat[s]
                return b"".join(out)[:size]
            return read_chain(start)[:size]
    raise KeyError(want)


def _rk(v: int) -> float:
    if v & 2:
        x = float(v >> 2 if not (v & 0x80000000) else (v >> 2) - (1 << 30))
    else:
        x = struct.unpack("<d", struct.pack("<Q", (v & 0xFFFFFFFC) << 32))[0]
    return x / 100.0 if v & 1 else x


def _records(wb: bytes):
    pos = 0
    while pos + 4 <= len(wb):
        rtype, rlen = struct.unpack_from("<HH", wb, pos)
        yield pos, rtype, wb[pos + 4:pos + 4 + rlen]
        pos += 4 + rlen


def _read_sst(wb: bytes, start: int):
    """Decode the shared string table (SST + CONTINUE records)."""
    chunks = []
    pos = start
    rtype, rlen = struct.unpack_from("<HH", wb, pos)
    chunks.append(wb[pos + 4:pos + 4 + rlen])
    pos += 4 + rlen
    while True:
        rtype, rlen = struct.unpack_from("<HH", wb, pos)
        if rtype != 0x003C:
            break
        chunks.append(wb[pos + 4:pos + 4 + rlen])
        pos += 4 + rlen
    first = chunks[0]
    n_unique = struct.unpack_from("<I", first, 4)[0]
    strings = []
    ci, off = 0, 8

    def need(n):
        nonlocal ci, off
        if off + n > len(chunks[ci]):
            ci += 1
            off = 0

    for _ in range(n_unique):
        need(3)
        c = chunks[ci]
        cch = struct.unpack_from("<H", c, off)[0]
        flags = c[off + 2]
        off += 3
        rich = flags & 0x08
        ext = flags & 0x04
        n_runs = ext_len = 0
        if rich:
            n_runs = struct.unpack_from("<H", chunks[ci], off)[0]
            off += 2
        if ext:
            ext_len = struct.unpack_from("<I", chunks[ci], off)[0]
            off += 4
        out = []
        remaining = cch
        wide = flags & 0x01
        while remaining > 0:
            c = chunks[ci]
            bpc = 2 if wide else 1
            avail = (len(c) - off) // bpc
            take = min(avail, remaining)
            raw = c[off:off + take * bpc]
            out.append(raw.decode("utf-16-le" if wide else "latin-1"))
            off += take * bpc
            remaining -= take
            if remaining > 0:
                ci += 1
                off = 0
                wide = chunks[ci][0] & 0x01
                off = 1
        skip = 4 * n_runs + ext_len
        while skip > 0:
            c = chunks[ci]
            step = min(skip, len(c) - off)
            off += step
            skip -= step
            if skip > 0:
                ci += 1
                off = 0
        strings.append("".join(out))
    return strings


def read_workbook(path: str):
    """Return {sheet name: {(row, col): value}} for numeric and label cells."""
    wb = _ole_stream(Path(path).read_bytes(), "Workbook")
    sheets, sst = [], []
    for pos, rtype, body in _records(wb):
        if rtype == 0x0085:  # BOUNDSHEET
            bof = struct.unpack_from("<I", body, 0)[0]
            nlen = body[6]
            wide = body[7] & 1
            name = body[8:8 + nlen * (2 if wide else 1)].decode("utf-16-le" if wide else "latin-1")
            sheets.append((name, bof))
        elif rtype == 0x00FC:
            sst = _read_sst(wb, pos)
    book = {}
    for name, bof in sheets:
        cells = {}
        for _, rtype, body in _records(wb[bof:]):
            if rtype == 0x000A:  # EOF
                break
            if rtype == 0x0203:  # NUMBER
                r, c = struct.unpack_from("<HH", body, 0)
                cells[(r, c)] = struct.unpack_from("<d", body, 6)[0]
            elif rtype == 0x027E:  # RK
                r, c = struct.unpack_from("<HH", body, 0)
                cells[(r, c)] = _rk(struct.unpack_from("<I", body, 6)[0])
            elif rtype == 0x00BD:  # MULRK
                r, c0 = struct.unpack_from("<HH", body, 0)
                n = (len(body) - 6) // 6
                for k in range(n):
                    cells[(r, c0 + k)] = _rk(struct.unpack_from("<I", body, 4 + 6 * k + 2)[0])
            elif rtype == 0x0006:  # FORMULA: cached result
                r, c = struct.unpack_from("<HH", body, 0)
                res = body[6:14]
                if res[6:8] != b"\xff\xff":
                    cells[(r, c)] = struct.unpack("<d", res)[0]
            elif rtype == 0x00FD:  # LABELSST
                r, c = struct.unpack_from("<HH", body, 0)
                cells[(r, c)] = sst[struct.unpack_from("<I", body, 6)[0]]
        book[name] = cells
    return book


def main(argv):
    path = argv[1] if len(argv) > 1 else DEFAULT_XLS
    book = read_workbook(path)
    out = {
        "source": "report/cva-bcn-accuracy-mean-squares_code_fixed (2).xls",
        "setup": {"scenarios": book["Resume"][(13, 2)], "expiry": 25.0, "dt": 0.01,
                  "notional": 10.0, "K": 0.067, "dtau": 0.5, "rng": "cuRAND PSEUDO_DEFAULT seed 1234"},
        "ee_rows": {},
        "resume": [],
    }
    row_names = {0: "cpu", 1: "single_gpu", 2: "multi_gpu"}
    for sheet, cells in book.items():
        if sheet == "Resume":
            continue
        rows = {}
        for r, label in row_names.items():
            rows[label] = [cells[(r, c)] for c in range(51)]
        out["ee_rows"][sheet] = rows
    header = [book["Resume"].get((0, c)) for c in range(12)]
    for r in range(1, 13):
        row = [book["Resume"].get((r, c)) for c in range(12)]
        if any(v is not None for v in row):
            out["resume"].append(dict(zip([h or "col%d" % i for i, h in enumerate(header)], row)))
    dst = Path(__file__).with_name("xls_ee_profiles.json")
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv)
