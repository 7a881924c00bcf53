// .conf loader: a file of "$title\n<body>" sections.
//
// Behaviour follows src/config.c:19-106: the file is read whole, a leading
// '$' (or the text before the first '$') starts the first title, each title
// runs to the end of its line and each body runs to the next '$'.  Delimiter
// runs collapse (strtok semantics), at most 8 sections are kept, and a second
// load_config without unload_config appends to the sections already loaded.
// Section pointers stay valid until unload_config.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_common.h"

namespace {

constexpr size_t kMaxSections = 8;   // src/config.c:12

struct ConfigState {
    std::vector<std::unique_ptr<char[]>> buffers;   // every loaded file
    std::vector<const char *> titles;
    std::vector<const char *> bodies;
};

ConfigState g_config;

// strtok_r with the delimiter-run semantics the .conf format relies on.
char *next_token(char *start, const char *delims, char **save) {
    return strtok_r(start, delims, save);
}

}  // namespace

extern "C" int load_config(const char *relative_path) {
    std::FILE *f = std::fopen(relative_path, "rb");
    if (!f) {
        note(2, "Could not read file %s\n", relative_path);
        return 1;
    }
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (len < 0) len = 0;
    std::unique_ptr<char[]> text(new char[static_cast<size_t>(len) + 1]);
    size_t got = std::fread(text.get(), 1, static_cast<size_t>(len), f);
    std::fclose(f);
    text[got] = '\0';

    char *buf = text.get();
    char *save = nullptr;
    char *title;
    if (got > 0 && buf[0] == '$') {
        title = next_token(buf + 1, "\n", &save);
    } else {
        next_token(buf, "$", &save);          // text before the first '$'
        title = next_token(nullptr, "\n", &save);
    }
    while (title != nullptr && g_config.titles.size() < kMaxSections) {
        char *body = next_token(nullptr, "$", &save);
        if (body != nullptr) {
            g_config.titles.push_back(title);
            g_config.bodies.push_back(body);
        }
        title = next_token(nullptr, "\n", &save);
    }
    g_config.buffers.push_back(std::move(text));
    return 0;
}

extern "C" const char *get_config_section(const char *heading) {
    for (size_t i = 0; i < g_config.titles.size(); ++i)
        if (std::strcmp(heading, g_config.titles[i]) == 0) return g_config.bodies[i];
    return nullptr;
}

extern "C" void unload_config(void) {
    g_config.titles.clear();
    g_config.bodies.clear();
    g_config.buffers.clear();
}
