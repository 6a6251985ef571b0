// Exercises include/sycldb.hpp, the C++ layer with the reference's operator
// interface, through the C ABI.  Built with the library (sycldb_amd/build.py).
//   api_check host    -- host-only checks (op parsing, exception mapping)
//   api_check device  -- a small pipeline on the GPU against values computed
//                        here: selection, semi join, full join, projection,
//                        sum, group-by, count; the allocating filter_join,
//                        full_join(TableData) and group_by_aggregate
//                        (ColumnData) overloads, count_true_flags_async and
//                        the fused bundle entry
// Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sycldb.hpp"

#define REQUIRE(c)                                                                                                     \
    do {                                                                                                               \
        if (!(c)) {                                                                                                    \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);                                       \
            return 1;                                                                                                  \
        }                                                                                                              \
    } while (0)

static int host_checks()
{
    using namespace sycldb;
    REQUIRE(get_comp_op("<") == SDB_LT);
    REQUIRE(get_comp_op("=") == SDB_EQ); // unknown -> EQ
    REQUIRE(get_logical_op("AND") == SDB_AND);
    REQUIRE(get_logical_op("") == SDB_NONE);
    REQUIRE(is_filter_logical("AND") && !is_filter_logical(">="));
    REQUIRE(get_op_from_string("*") == SDB_MUL);
    bool threw = false;
    try {
        get_op_from_string("%");
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    REQUIRE(threw);
    return 0;
}

template <class T> static T *dev(const std::vector<T> &h)
{
    void *p = nullptr;
    sycldb::check(sdb_device_alloc(&p, h.size() * sizeof(T)));
    sycldb::check(sdb_copy_to_device(p, h.data(), h.size() * sizeof(T), nullptr));
    return static_cast<T *>(p);
}

template <class T> static std::vector<T> host(const T *d, size_t n)
{
    std::vector<T> h(n);
    sycldb::check(sdb_copy_to_host(h.data(), d, n * sizeof(T), nullptr));
    return h;
}

struct NoAllocator {};

static int device_checks()
{
    using namespace sycldb;
    Queue q;
    const int n = 1000, dim = 50;
    std::vector<int> key(n), price(n), disc(n), dkey(dim), dval(dim);
    std::vector<unsigned char> one(n, 1), dflags(dim, 0);
    for (int i = 0; i < n; i++) {
        key[i] = 1 + (i * 37) % dim;
        price[i] = 1000 + (i * 7919) % 5000;
        disc[i] = i % 11;
    }
    for (int j = 0; j < dim; j++) {
        dkey[j] = j + 1;
        dval[j] = j % 5;
        dflags[j] = (j % 3) == 0;
    }
    int *d_key = dev(key), *d_price = dev(price), *d_disc = dev(disc), *d_dkey = dev(dkey), *d_dval = dev(dval);
    bool *d_flags = reinterpret_cast<bool *>(dev(one));
    bool *d_dflags = reinterpret_cast<bool *>(dev(dflags));
    bool *d_ht = reinterpret_cast<bool *>(dev(std::vector<unsigned char>(dim, 0)));
    int *d_kv = dev(std::vector<int>(2 * dim, 0));
    int *d_val = dev(std::vector<int>(n, 0));
    int *d_rev = dev(std::vector<int>(n, 0));
    uint64_t *d_sum = dev(std::vector<uint64_t>(1, 0));
    uint64_t *d_cnt = dev(std::vector<uint64_t>(1, 0));

    // discount in [1, 3] (two selections, the second AND-ed), then a semi
    // join with the dimension's flagged keys, a full join for its value,
    // revenue = price * discount, sum and count
    Event e1 = selection(d_flags, d_disc, ">=", 1, "", n, q, {});
    Event e2 = selection(d_flags, d_disc, "<=", 3, "AND", n, q, {e1});
    Event b1 = build_keys_ht(d_dkey, d_dflags, dim, d_ht, dim, 1, q, {});
    Event b2 = build_key_vals_ht(d_dkey, d_dval, d_dflags, dim, d_kv, dim, 1, q, {});
    Event j1 = filter_join(d_key, d_flags, n, d_ht, 1, dim, q, {e2, b1});
    Event j2 = full_join(d_key, d_val, d_flags, n, d_kv, 1, dim, q, {j1, b2});
    Event p1 = perform_operation(d_rev, d_price, d_disc, d_flags, n, "*", q, {j2});
    Event a1 = aggregate_operation(d_rev, d_flags, n, d_sum, q, {p1});
    NoAllocator na;
    count_true_flags(d_flags, n, q, na, d_cnt, {a1}).wait();

    // group by the joined value: SUM(revenue)
    const int *contents[1] = {d_val};
    int mins[1] = {0}, maxs[1] = {4};
    int *d_gkeys = dev(std::vector<int>(5, 0));
    int *results[1] = {d_gkeys};
    uint64_t *d_gsum = dev(std::vector<uint64_t>(5, 0));
    unsigned *d_seen = reinterpret_cast<unsigned *>(dev(std::vector<uint32_t>(5, 0)));
    group_by_aggregate(contents, d_rev, maxs, mins, d_flags, n, 1, results, d_gsum, d_seen, 5, "SUM", q, {}).wait();

    // expected on the host
    uint64_t want_sum = 0, want_cnt = 0;
    std::vector<uint64_t> want_g(5, 0);
    std::vector<unsigned> want_seen(5, 0);
    for (int i = 0; i < n; i++) {
        const bool pass = disc[i] >= 1 && disc[i] <= 3 && dflags[key[i] - 1];
        if (!pass) continue;
        const int rev = price[i] * disc[i];
        want_sum += (uint64_t)(int64_t)rev;
        want_cnt++;
        want_g[dval[key[i] - 1]] += (uint64_t)(int64_t)rev;
        want_seen[dval[key[i] - 1]] = 1;
    }
    REQUIRE(host(d_sum, 1)[0] == want_sum);
    REQUIRE(host(d_cnt, 1)[0] == want_cnt);
    REQUIRE(count_true_flags(d_flags, n, q) == want_cnt);
    REQUIRE(host(d_gsum, 5) == want_g);
    REQUIRE(host(d_seen, 5) == want_seen);
    const std::vector<int> gk = host(d_gkeys, 5);
    for (int g = 0; g < 5; g++)
        if (want_seen[g]) REQUIRE(gk[g] == g);
    bool threw = false;
    try {
        perform_operation(d_rev, d_price, d_disc, d_flags, n, "%", q, {});
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    REQUIRE(threw);
    REQUIRE(count_true_flags_async(d_flags, n, q, {}).get() == want_cnt);

    // the allocating overloads (kernels/join.hpp:210-254, 357-461,
    // kernels/aggregation.hpp:411-539) with a device allocator
    {
        DeviceAllocator alloc;
        bool *f2 = reinterpret_cast<bool *>(dev(one));
        filter_join(d_dkey, d_dflags, dim, dim, 1, d_key, f2, n, nullptr, alloc, q, {}).wait();
        const std::vector<unsigned char> hf2 = host(reinterpret_cast<unsigned char *>(f2), n);
        for (int i = 0; i < n; i++) REQUIRE(hf2[i] == dflags[key[i] - 1]);

        // probe table: one column (the foreign key), build table: key and value
        int *pkey = dev(key);
        ColumnData<int> pcols[1] = {{pkey, false, false, 1, dim}};
        ColumnData<int> bcols[2] = {{d_dkey, false, false, 1, dim}, {d_dval, false, false, 0, 4}};
        TableData<int> probe{pcols, 1, n, 1, f2, -1, nullptr, 0, 0, "lineorder", {{7, 0}}};
        TableData<int> build{bcols, 2, dim, 2, d_dflags, 1, nullptr, 0, 0, "dim", {{0, 0}, {1, 1}}};
        full_join(probe, build, 7, 0, alloc, q, {}).wait();
        REQUIRE(probe.column_indices.count(7) == 0 && probe.column_indices.at(1 + 1) == 0);
        REQUIRE(pcols[0].min_value == 0 && pcols[0].max_value == 4);
        const std::vector<int> joined = host(pkey, n);
        for (int i = 0; i < n; i++)
            if (hf2[i]) REQUIRE(joined[i] == dval[key[i] - 1]);

        // GROUP BY the joined value: SUM(price)
        auto g = group_by_aggregate(pcols, d_price, f2, 1, n, "SUM", alloc, q, {});
        std::get<4>(g).wait();
        REQUIRE(std::get<1>(g) == 5);
        std::vector<uint64_t> gs(5, 0);
        std::vector<unsigned char> gf(5, 0);
        for (int i = 0; i < n; i++)
            if (hf2[i]) {
                gs[dval[key[i] - 1]] += (uint64_t)(int64_t)price[i];
                gf[dval[key[i] - 1]] = 1;
            }
        REQUIRE(host(std::get<3>(g), 5) == gs);
        REQUIRE(host(reinterpret_cast<unsigned char *>(std::get<2>(g)), 5) == gf);
        const std::vector<int> gk = host(std::get<0>(g)[0], 5);
        for (int k = 0; k < 5; k++)
            if (gf[k]) REQUIRE(gk[k] == k);
        delete[] std::get<0>(g);
        check(sdb_device_free(pkey));
        check(sdb_device_free(f2));
    }

    // the fused bundle entry: discount in [1, 3], the semi join, price *
    // discount and its sum in one call
    {
        bool *f3 = reinterpret_cast<bool *>(dev(one));
        int *rev3 = dev(std::vector<int>(n, 0));
        uint64_t *sum3 = dev(std::vector<uint64_t>(1, 0));
        std::vector<sdb_op> ops(5);
        for (sdb_op &o : ops) o = sdb_op{};
        ops[0].kind = SDB_OP_SELECT_LIT, ops[0].cmp = SDB_GE, ops[0].logic = SDB_AND, ops[0].value = 1;
        ops[1].kind = SDB_OP_SELECT_LIT, ops[1].cmp = SDB_LE, ops[1].logic = SDB_AND, ops[1].value = 3;
        for (int k = 0; k < 2; k++) {
            ops[k].col = d_disc;
            ops[k].in_flags = detail::u8(f3);
            ops[k].out_flags = detail::u8(f3);
        }
        ops[2].kind = SDB_OP_PROBE_SEMI, ops[2].col = d_key, ops[2].ht = d_ht, ops[2].ht_min = 1, ops[2].ht_max = dim;
        ops[2].in_flags = detail::u8(f3), ops[2].out_flags = detail::u8(f3);
        ops[3].kind = SDB_OP_BINOP_CC, ops[3].binop = SDB_MUL, ops[3].col = d_price, ops[3].col2 = d_disc;
        ops[3].res = rev3, ops[3].in_flags = detail::u8(f3);
        ops[4].kind = SDB_OP_SUM, ops[4].col = rev3, ops[4].in_flags = detail::u8(f3), ops[4].acc = sum3;
        bundle_execute(ops, n, q).wait();
        REQUIRE(host(sum3, 1)[0] == want_sum);
        REQUIRE(count_true_flags(f3, n, q) == want_cnt);
        for (void *p : {(void *)f3, (void *)rev3, (void *)sum3}) check(sdb_device_free(p));
    }
    for (void *p : {(void *)d_key, (void *)d_price, (void *)d_disc, (void *)d_dkey, (void *)d_dval, (void *)d_flags,
                    (void *)d_dflags, (void *)d_ht, (void *)d_kv, (void *)d_val, (void *)d_rev, (void *)d_sum,
                    (void *)d_cnt, (void *)d_gkeys, (void *)d_gsum, (void *)d_seen})
        check(sdb_device_free(p));
    return 0;
}

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "host";
    int rc = 0;
    try {
        if (host_checks()) rc = 1;
        else if (mode == "device" && device_checks()) rc = 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        rc = 1;
    }
    // the bundle entry may have queued kernels with the background compiler:
    // it is stopped before the process exits
    sdb_codegen_shutdown();
    if (rc == 0) std::printf("ok\n");
    return rc;
}
