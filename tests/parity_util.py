"""Record-by-record comparison of device pipeline records with the oracle's
CPU pipeline (searchQueries3 -> alignQueries_cpu_mt -> chooseBestAlignment,
tie rule "first")."""

# device Records column -> oracle record field
FIELDS = (("read", "read"), ("strand", "strand"), ("ref_idx", "refIdx"), ("shift", "shift"),
          ("ref_pos", "refPos"), ("nbiggest", "nBiggest"), ("ref_distance", "refDistance"),
          ("score", "score"), ("start", "start"), ("end", "end"), ("alen", "len"),
          ("keep", "keep"), ("ntied", "nTied"))


def assert_records_match(got, want, ctx="", read0=0):
    """`got`: device Records; `want`: oracle pipeline records.  read0 is
    subtracted from the oracle's read indices (requests served alone number
    their reads from 0)."""
    assert got.n == len(want), (ctx, got.n, len(want))
    for i, w in enumerate(want):
        g = got.record(i)
        for a, b in FIELDS:
            wv = w[b] - read0 if b == "read" else w[b]
            assert g[a] == wv, (ctx, i, a, g[a], wv)
        assert g["qry_cons"] == w["qry_cons"] and g["ref_cons"] == w["ref_cons"], (ctx, i)
