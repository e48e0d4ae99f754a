// PrimaryVertexContext.h -- host snapshot of one work item's device state, with
// the accessors of the reference's context
// (include/ITSReconstruction/CA/PrimaryVertexContext.h:51-77).  The device
// state itself lives in HBM; a snapshot is downloaded on request
// (Tracker<true>::getPrimaryVertexContext).
#pragma once

#include <array>
#include <vector>

#include "Constants.h"
#include "Event.h"
#include "Road.h"

namespace o2 {
namespace ITS {
namespace CA {

class PrimaryVertexContext {
 public:
  static constexpr int TableSize = 401;

  const float3& getPrimaryVertex() const { return mPrimaryVertex; }
  std::array<std::vector<Cluster>, Constants::ITS::LayersNumber>& getClusters() { return mClusters; }
  std::array<std::vector<Cell>, Constants::ITS::CellsPerRoad>& getCells() { return mCells; }
  std::array<std::vector<int>, Constants::ITS::CellsPerRoad - 1>& getCellsLookupTable() { return mCellsLookupTable; }
  std::array<std::vector<std::vector<int>>, Constants::ITS::CellsPerRoad - 1>& getCellsNeighbours()
  {
    return mCellsNeighbours;
  }
  std::vector<Road>& getRoads() { return mRoads; }
  std::array<std::array<int, TableSize>, Constants::ITS::TrackletsPerRoad>& getIndexTables() { return mIndexTables; }
  std::array<std::vector<Tracklet>, Constants::ITS::TrackletsPerRoad>& getTracklets() { return mTracklets; }
  std::array<std::vector<int>, Constants::ITS::CellsPerRoad>& getTrackletsLookupTable() { return mTrackletsLookupTable; }

  float3 mPrimaryVertex{0, 0, 0};
  std::array<std::vector<Cluster>, Constants::ITS::LayersNumber> mClusters;
  std::array<std::vector<Cell>, Constants::ITS::CellsPerRoad> mCells;
  std::array<std::vector<int>, Constants::ITS::CellsPerRoad - 1> mCellsLookupTable;
  std::array<std::vector<std::vector<int>>, Constants::ITS::CellsPerRoad - 1> mCellsNeighbours;
  std::vector<Road> mRoads;
  std::array<std::array<int, TableSize>, Constants::ITS::TrackletsPerRoad> mIndexTables{};
  std::array<std::vector<Tracklet>, Constants::ITS::TrackletsPerRoad> mTracklets;
  std::array<std::vector<int>, Constants::ITS::CellsPerRoad> mTrackletsLookupTable;
};

}  // namespace CA
}  // namespace ITS
}  // namespace o2
