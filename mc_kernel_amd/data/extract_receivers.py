"""Writes gibraltar_receivers_200.txt: the latitude and longitude (degrees) of the 200
stations of the reference's Gibraltar benchmark (Benchmark_Gibraltar_event/receiver_200.dat,
station lines '<name> <lat> <lon> <nkernel>'), in file order.  Only the coordinates are
kept: the c4 workload defines its own kernels on them.  Run from the repository root with
the reference tree at /root/reference."""
import os
import sys

SRC = "/root/reference/Benchmark_Gibraltar_event/receiver_200.dat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gibraltar_receivers_200.txt")


def main():
    rows = []
    with open(SRC) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    i = 2
    while len(rows) < n:
        parts = lines[i].split()
        lat, lon, nk = float(parts[1]), float(parts[2]), int(parts[3])
        rows.append((lat, lon))
        i += 1 + nk
    with open(OUT, "w") as f:
        f.write("# lat lon (degrees) of the 200 stations of Benchmark_Gibraltar_event/receiver_200.dat\n")
        for lat, lon in rows:
            f.write("%.3f %.3f\n" % (lat, lon))
    print("%d receivers -> %s" % (len(rows), OUT), file=sys.stderr)


if __name__ == "__main__":
    main()
