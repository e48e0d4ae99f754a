// Road.h -- output records of the reference API: Tracklet, Cell, Road
// (include/ITSReconstruction/CA/{Tracklet,Cell,Road}.h; Road.cxx:24-50).
#pragma once

#include <array>

#include "Constants.h"
#include "Definitions.h"

namespace o2 {
namespace ITS {
namespace CA {

struct Tracklet {
  int firstClusterIndex;
  int secondClusterIndex;
  float tanLambda;
  float phiCoordinate;
};

class Cell {
 public:
  Cell(int c1, int c2, int c3, int t1, int t2, const float3& normal, float curvature, int level = 1)
      : mFirstClusterIndex{c1}, mSecondClusterIndex{c2}, mThirdClusterIndex{c3}, mFirstTrackletIndex{t1},
        mSecondTrackletIndex{t2}, mNormalVectorCoordinates(normal), mCurvature{curvature}, mLevel{level}
  {
  }
  int getFirstClusterIndex() const { return mFirstClusterIndex; }
  int getSecondClusterIndex() const { return mSecondClusterIndex; }
  int getThirdClusterIndex() const { return mThirdClusterIndex; }
  int getFirstTrackletIndex() const { return mFirstTrackletIndex; }
  int getSecondTrackletIndex() const { return mSecondTrackletIndex; }
  int getLevel() const { return mLevel; }
  float getCurvature() const { return mCurvature; }
  const float3& getNormalVectorCoordinates() const { return mNormalVectorCoordinates; }
  void setLevel(int level) { mLevel = level; }

 private:
  int mFirstClusterIndex;
  int mSecondClusterIndex;
  int mThirdClusterIndex;
  int mFirstTrackletIndex;
  int mSecondTrackletIndex;
  float3 mNormalVectorCoordinates;
  float mCurvature;
  int mLevel;
};

class Road {
 public:
  Road() : mCellIds{}, mRoadSize{0}, mLabel{Constants::ITS::UnusedIndex}, mIsFakeRoad{false} { resetRoad(); }
  Road(int cellLayer, int cellId) : Road() { addCell(cellLayer, cellId); }

  int getRoadSize() const { return mRoadSize; }
  int getLabel() const { return mLabel; }
  void setLabel(int label) { mLabel = label; }
  bool isFakeRoad() const { return mIsFakeRoad; }
  void setFakeRoad(bool fake) { mIsFakeRoad = fake; }
  int& operator[](const int& i) { return mCellIds[i]; }
  int operator[](const int& i) const { return mCellIds[i]; }

  void resetRoad()
  {
    mCellIds.fill(Constants::ITS::UnusedIndex);
    mRoadSize = 0;
  }
  void addCell(int cellLayer, int cellId)
  {
    if (mCellIds[cellLayer] == Constants::ITS::UnusedIndex) ++mRoadSize;
    mCellIds[cellLayer] = cellId;
  }

 private:
  std::array<int, Constants::ITS::CellsPerRoad> mCellIds;
  int mRoadSize;
  int mLabel;
  bool mIsFakeRoad;
};

}  // namespace CA
}  // namespace ITS
}  // namespace o2
