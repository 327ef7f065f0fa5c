"""Writes tests/golden/mobydick_excerpt.txt: the first lines (about 48 KB) of
the text the reference's BWT test sorts (bwt/src/mobydick.txt, read by
bwt/src/bwttest/bwttest.cpp:9,33-44).  The text is a data file of that test,
copied unchanged up to a line boundary; run here, where the reference is
present, and commit the output."""
import os

SRC = "/root/reference/bwt/src/mobydick.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mobydick_excerpt.txt")
LIMIT = 48 * 1024


def main():
    with open(SRC, "rb") as f:
        data = f.read()
    cut = data.rfind(b"\n", 0, LIMIT) + 1
    with open(OUT, "wb") as f:
        f.write(data[:cut])
    print(f"{OUT}: {cut} bytes")


if __name__ == "__main__":
    main()
