// The C ABI's opaque handles (include/sycldb.h), shared by the files that
// implement their functions.
#pragma once

#include "engine.hpp"

struct sdb_engine {
    sdb::Engine eng;
};

struct sdb_result {
    sdb::ResultTable t;
};
