"""Host logic of the bucket-table autotuner (fastjet_gpu_amd/autotune.py): the
candidate tables of its whole-batch stage, built from per-interval winners.
No GPU: these are the pure table-building functions."""
from fastjet_gpu_amd import autotune as at

# per-interval winners as the first stage returns them: (max_n, (engine, threads, slots, variant))
CHOICES = [(32, (1, 64, 8, 0)), (64, (1, 64, 1, 0)), (128, (1, 64, 2, 0)), (256, (1, 64, 4, 0)),
           (1536, (2, 64, 5, 3)), (4096, (2, 64, 4, 1)), (20000, (2, 64, 4, 0))]


def test_merge_joins_neighbours_and_opens_the_last_row():
    rows = at.merge(CHOICES)
    assert rows == [(32, 1, 64, 8, 0), (64, 1, 64, 1, 0), (128, 1, 64, 2, 0), (256, 1, 64, 4, 0),
                    (1536, 2, 64, 5, 3), (4096, 2, 64, 4, 1), (at.LAST_ROW, 2, 64, 4, 0)]
    # equal neighbours merge into one row; a dense last row gets a tiled row after it
    assert at.merge([(64, (1, 64, 1, 0)), (128, (1, 64, 1, 0))]) == [(128, 1, 64, 1, 0), (at.LAST_ROW, 2, 64, 0, 0)]


def test_tiled_alone_one_row_per_variant():
    configs = [(128, 1, 64, 2, 0), (at.LAST_ROW, 2, 64, 0, 0), (at.LAST_ROW, 2, 64, 0, 1), (at.LAST_ROW, 2, 64, 0, 3)]
    tables = at.tiled_alone(configs)
    assert [rows for _, rows in tables] == [[(at.LAST_ROW, 2, 64, 4, v)] for v in (0, 1, 3)]


def test_simple_tables_put_dense_rows_before_the_given_tiled_row():
    last = (2, 64, 4, 1)
    tables = dict(at.simple_tables(CHOICES, last))
    assert tables["dense<= 128 + tiled"] == [(32, 1, 64, 8, 0), (64, 1, 64, 1, 0), (128, 1, 64, 2, 0),
                                             (at.LAST_ROW, 2, 64, 4, 1)]
    assert tables["one dense row <= 256 + tiled"] == [(256, 1, 64, 4, 0), (at.LAST_ROW, 2, 64, 4, 1)]
    # no cut without a dense winner ending exactly there
    assert not any(name.startswith("dense<= 512") for name in tables)
    for rows in tables.values():  # every candidate table is sorted and ends with an open tiled row
        assert [r[0] for r in rows] == sorted(r[0] for r in rows)
        assert rows[-1][0] == at.LAST_ROW and rows[-1][1] == 2


def test_written_table_runs_its_last_row_on_eight_waves():
    rows = at.coop_top_row([(4096, 2, 64, 4, 3), (at.LAST_ROW, 2, 64, 4, 1)])
    assert rows == [(4096, 2, 64, 4, 3), (at.LAST_ROW, 2, 512, 4, 1)]
    # a dense last row is left alone (merge always ends with a tiled row)
    assert at.coop_top_row([(at.LAST_ROW, 1, 64, 4, 0)]) == [(at.LAST_ROW, 1, 64, 4, 0)]
