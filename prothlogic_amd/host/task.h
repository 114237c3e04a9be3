// task.h -- the task record of P/task.h (a label plus string payload) that
// the worker entry points take.
#pragma once

#include <cstdint>
#include <map>
#include <string>

enum ProveMode {
    PROVE_PROTH_FF = 1,
};

struct Task {
    unsigned int n;
    uint64_t k;
    std::string label;
    std::map<std::string, std::string> payload;
};
