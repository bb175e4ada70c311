/*
 * video.h -- minimal planar video container with the reference's layout
 * and accessors (VideoSize / Video<T>, src/Utilities/LibVideoT.hpp:40-190):
 * index = t*w*h*c + c*w*h + y*w + x, 32-bit unsigned.
 */
#ifndef VBM3D_AMD_VIDEO_H
#define VBM3D_AMD_VIDEO_H

#include <vector>

struct VideoSize
{
	unsigned width = 0, height = 0, frames = 0, channels = 0;
	unsigned wh = 0, whc = 0, whcf = 0, whf = 0;

	VideoSize() {}
	VideoSize(unsigned w, unsigned h, unsigned f, unsigned c)
		: width(w), height(h), frames(f), channels(c) { update_fields(); }

	void update_fields()
	{
		wh = width * height;
		whc = wh * channels;
		whcf = whc * frames;
		whf = wh * frames;
	}
	bool operator==(const VideoSize& o) const
	{
		return width == o.width && height == o.height && channels == o.channels && frames == o.frames;
	}
	bool operator!=(const VideoSize& o) const { return !(*this == o); }
	unsigned index(unsigned x, unsigned y, unsigned t, unsigned c) const
	{
		return t * whc + c * wh + y * width + x;
	}
	void coords(unsigned idx, unsigned& x, unsigned& y, unsigned& t, unsigned& c) const
	{
		t = idx / whc;
		c = (idx % whc) / wh;
		y = (idx % wh) / width;
		x = idx % width;
	}
};

template <class T>
class Video
{
public:
	VideoSize sz;
	std::vector<T> data;

	Video() {}
	Video(unsigned w, unsigned h, unsigned f, unsigned c = 1) : sz(w, h, f, c), data(sz.whcf) {}
	Video(unsigned w, unsigned h, unsigned f, unsigned c, T val) : sz(w, h, f, c), data(sz.whcf, val) {}
	explicit Video(const VideoSize& s) : sz(s), data(s.whcf) {}
	Video(const VideoSize& s, T val) : sz(s), data(s.whcf, val) {}

	void clear() { sz = VideoSize(); data.clear(); }
	void resize(unsigned w, unsigned h, unsigned f, unsigned c = 1) { resize(VideoSize(w, h, f, c)); }
	void resize(const VideoSize& s)
	{
		if (s != sz) { sz = s; sz.update_fields(); data.assign(sz.whcf, T(0)); }
	}
	T& operator()(unsigned idx) { return data[idx]; }
	T operator()(unsigned idx) const { return data[idx]; }
	T& operator()(unsigned x, unsigned y, unsigned t, unsigned c = 0) { return data[sz.index(x, y, t, c)]; }
	T operator()(unsigned x, unsigned y, unsigned t, unsigned c = 0) const { return data[sz.index(x, y, t, c)]; }
};

#endif
