"""The reference's text state format written out from the stream operators it
uses, as an independent second implementation for the tests:
PathIntegralMonteCarlo::saveState (src/pimc.cpp:926-986) with
operator<<(ostream, std::array) and operator<<(ostream, DynamicArray<T,2>)
(include/common.h:244-272) under std::setprecision(16), and MTRand::save words
(include/MersenneTwister.h:370-377)."""


def _array(rows, cols, items):
    out = [f"(0,{rows - 1}) x (0,{cols - 1})\n[ "]
    for i in range(rows):
        for j in range(cols):
            out.append(items(i, j) + " ")
        if i < rows - 1:
            out.append("\n  ")
    out.append("]\n")
    return "".join(out)


def state_text(num_worldlines, counters, beads, nxt, prv, worm, random_words):
    """beads [M][N][3], nxt/prv [M][N][2] (XXX = -1), worm [M][N], counters
    [(a, b), ...] (totals first), random_words: 625 ints or None."""
    M, N = len(worm), len(worm[0])
    s = [f"{num_worldlines}\n"]
    for a, b in counters:
        s.append("%16d\t%16d\n" % (a, b))
    # each array is followed by the std::endl saveState adds
    s.append(_array(M, N, lambda i, j: "(" + ",".join("%.16g" % x for x in beads[i][j]) + ")") + "\n")
    s.append(_array(M, N, lambda i, j: "(%d,%d)" % tuple(nxt[i][j])) + "\n")
    s.append(_array(M, N, lambda i, j: "(%d,%d)" % tuple(prv[i][j])) + "\n")
    s.append(_array(M, N, lambda i, j: "%d" % worm[i][j]) + "\n")
    if random_words is not None:
        s.append("".join("%d " % w for w in random_words))
    s.append("\n")
    return "".join(s)
