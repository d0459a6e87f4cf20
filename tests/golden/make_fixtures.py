"""Extract small golden fixtures from the reference's captured outputs.

Run in a container where the reference tree is mounted (default
/root/reference); the fixtures it writes under tests/golden/ are committed so
that the tests never need the reference at run time.

Sources (data files the reference holds; no reference code is run):
  3.txt  lines 1-65536: forward FFT of generate_data(256, 256, seed=42), first
         128 transforms, interleaved re/im, one value per line "a.....b"
         (both columns equal); line 65538 is a timing line.
  2.txt  lines 1-131072: "mcfft ... fftw" for generate_data(262144, 1, 42);
         the second column is FFTW FP64, first 65536 complex outputs; the
         first column is the reference's (wrong) output; line 131073 holds the
         reference's accuracy metric for that run.
  1.log  lines 2-513: gen_rev(512, {16,16,2}) printed by setup();
         line 515 the accuracy metric of that run.
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def parse_3txt():
    vals = []
    with open(os.path.join(REF, "3.txt")) as f:
        for i, line in enumerate(f):
            if i >= 65536:
                break
            a, b = line.strip().split(".....")
            vals.append((float(a), float(b)))
    v = np.array(vals)
    assert np.array_equal(v[:, 0], v[:, 1])
    return np.ascontiguousarray(v[:, 0]).view(np.complex128)  # 32768 complex = 128 transforms of 256


def parse_2txt():
    a_col, b_col = [], []
    metric = None
    with open(os.path.join(REF, "2.txt")) as f:
        for i, line in enumerate(f):
            line = line.strip()
            if " ... " in line:
                a, b = line.split(" ... ")
                a_col.append(float(a))
                b_col.append(float(b))
            elif line:
                metric = float(line)
    return np.array(a_col).view(np.complex128), np.array(b_col).view(np.complex128), metric


def parse_1log():
    ints, metric = [], None
    with open(os.path.join(REF, "1.log")) as f:
        lines = [l.strip() for l in f]
    inside = False
    for l in lines:
        if l.startswith("^^^^"):
            inside = not inside
            continue
        if inside:
            ints.append(int(l))
        elif l:
            metric = float(l)
    return np.array(ints, dtype=np.int32), metric


def main():
    rng = np.random.default_rng(12345)
    t3 = parse_3txt()
    idx3 = np.sort(rng.choice(np.arange(2048, t3.size), 2048, replace=False))
    np.savez(os.path.join(OUT, "ref_3txt_n256_b256_seed42.npz"),
             head=t3[:2048], sample_idx=idx3, sample_val=t3[idx3],
             n=256, n_batch=256, seed=42, n_dumped=t3.size)
    mc, fftw, metric2 = parse_2txt()
    idx2 = np.sort(rng.choice(np.arange(2048, fftw.size), 2048, replace=False))
    np.savez(os.path.join(OUT, "ref_2txt_n262144_b1_seed42.npz"),
             fftw_head=fftw[:2048], fftw_sample_idx=idx2, fftw_sample_val=fftw[idx2],
             mcfft_head=mc[:2048], metric=metric2, n=262144, n_batch=1, seed=42, n_dumped=fftw.size)
    rev, metric1 = parse_1log()
    np.savez(os.path.join(OUT, "ref_1log_genrev512.npz"), rev=rev, radices=np.array([16, 16, 2]), metric=metric1)
    print("wrote fixtures:", t3.size, fftw.size, rev.size, metric1, metric2)


if __name__ == "__main__":
    main()
