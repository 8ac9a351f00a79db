// hierarchy.h -- drop-in for the reference's src/hierarchy.h (:1-29): the
// taxonomy store (taxonomy_index.tsv + sequences_by_category/<cat>.fasta).
#pragma once
#include <filesystem>
#include <string>
#include <unordered_map>
#include <vector>

class Hierarchy {
public:
    Hierarchy(std::string seq_dir, std::string taxes);

    const std::vector<std::string>& get_sequences(std::string category);

    void append_sequence(std::string category, std::string seq);

    std::vector<std::string> get_order(std::string order);

    const std::vector<std::string>* get_children(std::string category);

private:
    std::string make_seq_id(const std::string& category);

    std::filesystem::path seq_dir;
    std::string taxes;
    std::unordered_map<std::string, std::vector<std::string>> children;
    std::unordered_map<std::string, std::vector<std::string>> seq_cache;
};
