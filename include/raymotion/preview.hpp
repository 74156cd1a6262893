#pragma once
// Live-preview frame channel (interface of include/raymotion/preview.hpp in the
// reference): frames are written to descriptor 3 as "RAY1 w h frame size\n"
// followed by the RGB bytes; one byte on stdin grants the next frame.
#include <cstdio>
#include <unistd.h>

#include "raymotion/runtime.hpp"

namespace raymotion {

class PreviewStream {
    FILE* stream;

public:
    PreviewStream() : stream(fdopen(dup(3), "wb")) {
        if (!stream) throw std::runtime_error("preview frame channel unavailable");
    }
    ~PreviewStream() {
        if (stream) std::fclose(stream);
    }
    PreviewStream(const PreviewStream&) = delete;
    PreviewStream& operator=(const PreviewStream&) = delete;
    FrameSink sink() {
        return [this](const std::vector<uint8_t>& rgb, int w, int h, int frame) {
            if (std::fprintf(stream, "RAY1 %d %d %d %zu\n", w, h, frame, rgb.size()) < 0 ||
                std::fwrite(rgb.data(), 1, rgb.size(), stream) != rgb.size() || std::fflush(stream) != 0)
                throw std::runtime_error("preview channel closed");
            if (std::getchar() == EOF) throw RenderComplete{};
        };
    }
};

}  // namespace raymotion
