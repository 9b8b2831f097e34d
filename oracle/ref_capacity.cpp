// Driver for the reference's own partition_capacity() (header compiled from
// where it lies in the reference tree; see oracle/ref.mk).  Reads
// "total parts imbalance" triples on stdin, prints the capacity or "error".
#include "partition_capacity.hpp"

#include <iostream>

int main() {
    unsigned long long total;
    int parts;
    double imbalance;
    while (std::cin >> total >> parts >> imbalance) {
        try {
            std::cout << gpart::partition_capacity(total, parts, imbalance) << '\n';
        } catch (const std::exception&) {
            std::cout << "error\n";
        }
    }
    return 0;
}
