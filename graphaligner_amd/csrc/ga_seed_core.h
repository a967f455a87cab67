// Per-read seed lookup and reverse complement, shared by the gfx950 kernels and
// the host test build (same source, so the CPU suite checks the kernel logic).
#pragma once
#include <stdint.h>
#include "ga_device_types.h"

#if defined(__HIPCC__)
#define GA_SEED_HD __host__ __device__ inline
#else
#define GA_SEED_HD inline
#endif

namespace gad {

GA_SEED_HD int readBaseCode(unsigned char c)
{
	switch (c)
	{
		case 'a': case 'A': return 0;
		case 'c': case 'C': return 1;
		case 'g': case 'G': return 2;
		case 't': case 'T': return 3;
	}
	return -1;
}

// The seeder's 64-bit k-mer hash (src/MinimizerSeeder.cpp:45-54).
GA_SEED_HD uint64_t seedKmerHash(uint64_t key)
{
	key = (~key) + (key << 21);
	key = key ^ (key >> 24);
	key = (key + (key << 3)) + (key << 8);
	key = key ^ (key >> 14);
	key = (key + (key << 2)) + (key << 4);
	key = key ^ (key >> 28);
	key = key + (key << 31);
	return key;
}

// Looks up every read k-mer that the seeder's iterateKmers reports
// (src/MinimizerSeeder.cpp:60-102: k-mers are restarted after a non-ACGT letter,
// and a repeat of the previous k-mer is reported again only once it lies a full
// window of w - k + 1 positions later). K-mers present in the table with fewer
// than maxCount positions are written to out in read order; returns their number.
GA_SEED_HD uint32_t lookupReadKmers(const char* s, uint64_t len, const SeedTableSlot* table, uint64_t mask, uint32_t k, uint32_t w, uint64_t maxCount, SeedMatch* out)
{
	uint32_t n = 0;
	if (len < k) return 0;
	const uint64_t realWindow = w - k + 1;
	const uint64_t kmask = ~(~0ull << (2 * k));
	uint64_t offset = 0;
	while (true)
	{
		while (offset < len && readBaseCode((unsigned char)s[offset]) < 0) offset++;
		if (offset + k > len) break;
		uint64_t kmer = 0;
		uint64_t bad = len;
		for (uint64_t i = 0; i < k; i++)
		{
			int c = readBaseCode((unsigned char)s[offset + i]);
			if (c < 0)
			{
				bad = i;
				break;
			}
			kmer = (kmer << 2) | (uint64_t)c;
		}
		if (bad != len)
		{
			offset += bad;
			continue;
		}
		uint64_t lastKmer = 0, lastPos = 0;
		bool restart = false;
		for (uint64_t i = k - 1; offset + i < len; i++)
		{
			if (i >= k)
			{
				int c = readBaseCode((unsigned char)s[offset + i]);
				if (c < 0)
				{
					offset += i;
					restart = true;
					break;
				}
				kmer = ((kmer << 2) & kmask) | (uint64_t)c;
			}
			if (i == k - 1 || lastKmer != kmer || lastPos <= offset + i - realWindow)
			{
				lastKmer = kmer;
				lastPos = offset + i;
				uint64_t slot = seedKmerHash(kmer) & mask;
				while (true)
				{
					uint64_t key = table[slot].kmer;
					if (key == kmer)
					{
						uint32_t c = table[slot].count;
						if (c < maxCount)
						{
							SeedMatch m;
							m.pos = (uint32_t)(offset + i);
							m.start = table[slot].start;
							m.count = c;
							out[n++] = m;
						}
						break;
					}
					if (key == ~0ull) break;
					slot = (slot + 1) & mask;
				}
			}
		}
		if (!restart) break;
	}
	return n;
}

// CommonUtils::Complement (src/CommonUtils.cpp:67-97); letters outside the
// table are kept as they are (the host drops such reads before they are used).
GA_SEED_HD char complementByte(char c)
{
	switch (c)
	{
		case 'A': case 'a': return 'T';
		case 'C': case 'c': return 'G';
		case 'T': case 't': return 'A';
		case 'G': case 'g': return 'C';
		case 'N': case 'n': return 'N';
		case 'U': case 'u': return 'A';
		case 'R': case 'r': return 'Y';
		case 'Y': case 'y': return 'R';
		case 'K': case 'k': return 'M';
		case 'M': case 'm': return 'K';
		case 'S': case 's': return 'S';
		case 'W': case 'w': return 'W';
		case 'B': case 'b': return 'V';
		case 'V': case 'v': return 'B';
		case 'D': case 'd': return 'H';
		case 'H': case 'h': return 'D';
	}
	return c;
}

} // namespace gad
