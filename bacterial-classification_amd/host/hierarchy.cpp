// hierarchy.cpp -- the taxonomy store, with the reference's file formats and
// behaviour (src/hierarchy.cpp:11-130): the first TSV line is skipped as a
// header; a rank token is kept when it holds "__" and does not end with it;
// categories map ' ' to '_'; a missing FASTA is an empty category; appended
// records are named <category>_appended_<n>.  The containers are the
// reference's (unordered_set per parent, unordered_map) because get_order()
// and get_children() expose their iteration order to the callers.
#include "hierarchy.h"

#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

namespace {
std::string spaces_to_underscores(std::string s) {
    for (char& c : s)
        if (c == ' ') c = '_';
    return s;
}
}  // namespace

Hierarchy::Hierarchy(std::string seq_dir_, std::string taxes_) : seq_dir(seq_dir_), taxes(taxes_) {
    std::unordered_map<std::string, std::unordered_set<std::string>> kids;
    std::ifstream in(taxes);
    std::string line;
    std::getline(in, line);  // header
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::string id, leaf, tax;
        std::stringstream fields(line);
        std::getline(fields, id, '\t');
        std::getline(fields, leaf, '\t');
        std::getline(fields, tax);
        std::vector<std::string> ranks;
        std::stringstream ts(tax);
        std::string tok;
        while (std::getline(ts, tok, ';')) {
            tok.erase(0, tok.find_first_not_of(' '));
            const bool named = tok.find("__") != std::string::npos &&
                               !(tok.size() >= 2 && tok.compare(tok.size() - 2, 2, "__") == 0);
            if (named) ranks.push_back(spaces_to_underscores(tok));
        }
        for (size_t i = 1; i < ranks.size(); ++i) kids[ranks[i - 1]].insert(ranks[i]);
    }
    for (auto& [parent, set] : kids) children[parent] = std::vector<std::string>(set.begin(), set.end());
}

const std::vector<std::string>& Hierarchy::get_sequences(std::string category) {
    category = spaces_to_underscores(category);
    auto it = seq_cache.find(category);
    if (it != seq_cache.end()) return it->second;
    std::vector<std::string> seqs;
    const std::filesystem::path fasta = seq_dir / (category + ".fasta");
    if (std::filesystem::exists(fasta)) {
        std::ifstream in(fasta);
        std::string line, cur;
        while (std::getline(in, line)) {
            if (line.rfind(">", 0) == 0) {
                if (!cur.empty()) {
                    seqs.push_back(cur);
                    cur.clear();
                }
            } else {
                cur += line;
            }
        }
        if (!cur.empty()) seqs.push_back(cur);
    }
    return seq_cache.emplace(category, std::move(seqs)).first->second;
}

void Hierarchy::append_sequence(std::string category, std::string seq) {
    category = spaces_to_underscores(category);
    auto it = seq_cache.find(category);
    if (it != seq_cache.end()) it->second.push_back(seq);
    const std::filesystem::path fasta = seq_dir / (category + ".fasta");
    std::filesystem::create_directories(fasta.parent_path());
    std::ofstream out(fasta, std::ios::app);
    if (!out) throw std::runtime_error("Failed to open FASTA for append: " + fasta.string());
    out << ">" << make_seq_id(category) << "\n" << seq << "\n";
}

std::vector<std::string> Hierarchy::get_order(std::string order) {
    std::vector<std::string> out;
    for (const auto& [parent, kids] : children)
        if (parent.rfind(order, 0) == 0) out.push_back(parent);
    return out;
}

const std::vector<std::string>* Hierarchy::get_children(std::string category) {
    auto it = children.find(category);
    return it == children.end() ? nullptr : &it->second;
}

std::string Hierarchy::make_seq_id(const std::string& category) {
    const std::string norm = spaces_to_underscores(category);
    auto it = seq_cache.find(norm);
    const size_t n = it == seq_cache.end() ? 0 : it->second.size();
    return norm + "_appended_" + std::to_string(n);
}
