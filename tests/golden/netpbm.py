"""Minimal binary Netpbm (P5 / P6) reader for the fixture generator."""
import numpy as np


def read(path):
    with open(path, "rb") as f:
        data = f.read()
    tokens, pos = [], 0
    while len(tokens) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            while data[pos:pos + 1] not in (b"\n", b""):
                pos += 1
            continue
        start = pos
        while not data[pos:pos + 1].isspace():
            pos += 1
        tokens.append(data[start:pos])
    magic, w, h, maxval = tokens[0], int(tokens[1]), int(tokens[2]), int(tokens[3])
    pos += 1  # single whitespace after maxval
    ch = 3 if magic == b"P6" else 1
    if magic not in (b"P5", b"P6") or maxval > 255:
        raise ValueError("only 8-bit P5/P6")
    arr = np.frombuffer(data, np.uint8, w * h * ch, pos)
    return (arr.reshape(h, w, ch) if ch == 3 else arr.reshape(h, w)), maxval
