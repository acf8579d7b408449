"""The /nearest request layer (svdb_nearest_request) without a GPU: the
validation the reference's handler does before it searches, in its order
and with its texts and statuses (compare_handler.c:287-386), plus the
checks of ?number= and ?metric=.  Every case here is answered before the
store is touched, so db may be NULL."""
import ctypes as C
import json

import pytest

import svdb


def call(body, dim=3, number=None, metric=None):
    L = svdb.lib()
    out = C.c_char_p()
    n = C.c_size_t(0)
    raw = body.encode() if isinstance(body, str) else body
    st = L.svdb_nearest_request(None, None, dim, raw, len(raw),
                                number.encode() if number is not None else None,
                                metric.encode() if metric is not None else None,
                                svdb.PRECISIONS["auto"], C.byref(out), C.byref(n))
    text = C.string_at(out, n.value).decode()
    svdb.free_response(out)
    return st, text


@pytest.mark.parametrize("body,want", [
    ("", (400, '{"error": "Empty data"}')),
    ("[1,2,", (400, '{"error": "Invalid JSON"}')),
    ("[1,2,3]]", (500, '{"error": "Database not initialised"}')),  # cJSON ignores what follows
    ("{\"a\":1,\"b\":2}", (400, '{"error": "Vector size mismatch"}')),
    ("[1,2]", (400, '{"error": "Vector size mismatch"}')),
    ("7", (400, '{"error": "Vector size mismatch"}')),
    ("[1,\"x\",3]", (400, '{"error": "Invalid vector data"}')),
    ("[1,null,3]", (400, '{"error": "Invalid vector data"}')),
    ("[[1,2,3],[4,5]]", (400, '{"error": "Vector size mismatch"}')),
    ("[[1,2,3],7]", (400, '{"error": "Invalid vector data"}')),
    ("[[1,2,3],[4,true,6]]", (400, '{"error": "Invalid vector data"}')),
    ("[01, 1., -2e0]", (500, '{"error": "Database not initialised"}')),  # numbers as cJSON reads them
    ("[1.2.3, 4, 5]", (400, '{"error": "Invalid JSON"}')),
    ("[+1, 2, 3]", (400, '{"error": "Invalid JSON"}')),
])
def test_validation_order_and_texts(body, want):
    assert call(body) == want


def test_object_members_count_as_elements():
    # cJSON_GetArraySize / GetArrayItem walk an object's members, so the
    # reference accepts {"a":1,"b":2,"c":3} as the vector [1, 2, 3]
    assert call('{"a":1,"b":2,"c":3}')[0] == 500  # passed validation; no database


@pytest.mark.parametrize("number,ok", [(None, True), ("1", True), ("128", True), ("0", False),
                                       ("129", False), ("-1", False), ("2x", False), ("", False)])
def test_number_parameter(number, ok):
    st, text = call("[1,2,3]", number=number)
    assert (st, text) == ((500, '{"error": "Database not initialised"}') if ok else
                          (400, '{"error": "Invalid number"}'))


@pytest.mark.parametrize("metric,ok", [("euclidean", True), ("cosine_similarity", True),
                                       ("dot", True), ("dot_product", True), ("manhattan", False)])
def test_metric_parameter(metric, ok):
    st, text = call("[1,2,3]", metric=metric)
    assert st == (500 if ok else 400)
    if not ok:
        assert json.loads(text) == {"error": "Unknown metric"}


def test_deep_nesting_is_invalid_not_a_crash():
    assert call("[" * 100000)[0] == 400
